"""Bond Monte Carlo gates of the reference, run on the HIP path.

The bond kernels are bit-exact against the oracle's sequential linearization (test_gpu_parity.py,
test_gpu_campaign_scale.py); these gates pin that shared restatement against the reference's own
measured chemistry, so a mistake common to both (a nucleation tip on the wrong carbon, swapped
attempt and accept draws, a walk over j < i, a missed valence cap) fails here.

- tests/catalyst-turnover.test.ts:51-146: box 20, seed 19, kT 1.1, start {C 779, O 3117, H 779,
  M 52, W 695}, 8 x 15,000 steps. The reference measured, at 120,000 steps, 117 nucleated chains,
  133 bare carbons and 56 amphiphiles with the tip taken on `j` (its historical bug), and 192, 33 and
  129 with the tip on `i`; the assertions sit between the two arms: chains >= 150, bare carbons <= 80,
  amphiphiles >= 100, at least one occupied centre, an unchanged census and a finite state.
- tests/soup-valence.test.ts:20-84: box 30, seed 11, kT 1.1, the default start of data/soup.json,
  45,000 steps: bonds formed, no carbon with more than 2 C-C or 1 C-O bonds, no head with more than
  headPlacement.chainCapacity bonds. The reference runs this with its default clay platelet; the
  platelet is not part of this build, so the run is clay-free.
"""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.analysis import findAmphiphiles
from protocell_soup.rules import loadSoup

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def _kinds(pos):
    return np.rint(pos.reshape(-1, 4)[:, 3]).astype(np.int64)


def _turnover_sample(s, soup, step):
    """The reference's sample(step): bond events, occupied catalyst centres, carbon-bearing bonded
    components (the cumulative nucleation count), amphiphiles and bare carbons."""
    pos = s.particles()
    slots = s.bondSlots().reshape(-1, 3).astype(np.int64)
    links = s.centerLinks()
    ev = s.events()
    kinds = _kinds(pos)
    kind_of = [m["kind"] for m in soup["monomers"]]
    is_c = np.array([kind_of[k] == "carbon" for k in kinds])
    is_m = np.array([kind_of[k] == "catalyst" for k in kinds])
    n = kinds.size
    deg = (slots != EMPTY).sum(axis=1)
    occupied = int((is_m & (links != EMPTY)).sum())
    bare_c = int((is_c & (deg == 0)).sum())
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i in range(n):
        for j in slots[i]:
            if j != EMPTY:
                a, b = find(i), find(int(j))
                if a != b:
                    parent[a] = b
    chains = len({find(i) for i in range(n) if is_c[i] and deg[i] > 0})
    amph = len(findAmphiphiles(pos, s.bonds(), soup["monomers"]))
    return {"step": step, "cc": ev.get("cc_bond", 0), "co": ev.get("co_bond", 0), "occupied": occupied, "chains": chains,
            "amph": amph, "bareC": bare_c, "desorb": s.desorbEvents()}


def test_catalyst_turns_over():
    soup = loadSoup()
    start = {"C": 779, "O": 3117, "H": 779, "M": 52, "W": 695}
    s = ps.createSoup({"box": [20, 20, 20], "seed": 19, "kT": 1.1, "clay": False, "start": start})
    try:
        trace = []
        for k in range(8):
            s.step(15000)
            trace.append(_turnover_sample(s, soup, (k + 1) * 15000))
            t = trace[-1]
            print(f"TURNOVER step={t['step']} cc_bond={t['cc']} co_bond={t['co']} occupiedCentres={t['occupied']}/52 "
                  f"chains={t['chains']} amphiphiles={t['amph']} bareCarbons={t['bareC']} "
                  f"desorbTimeout={t['desorb']['timeout']}", flush=True)
        nf = s.nonFiniteCount()
        inv = s.invariants()
    finally:
        s.dispose()
    last = trace[-1]
    assert nf == {"pos": 0, "vel": 0}
    assert inv["monomers"] == {"C": 779, "O": 3117, "H": 779, "M": 52, "W": 695, "K": 0}
    assert last["occupied"] > 0
    assert last["chains"] >= 150
    assert last["bareC"] <= 80
    assert last["amph"] >= 100


def test_valence_caps_hold_over_45000_steps():
    soup = loadSoup()
    kind_index = {m["id"]: k for k, m in enumerate(soup["monomers"])}
    carbon = kind_index[next(m["id"] for m in soup["monomers"] if m["kind"] == "carbon")]
    head = kind_index[next(m["id"] for m in soup["monomers"] if m["kind"] == "head")]
    s = ps.createSoup({"box": [30, 30, 30], "seed": 11, "kT": 1.1, "clay": False})
    try:
        s.step(45000)
        pos = s.particles()
        edges = s.bonds().reshape(-1, 2).astype(np.int64)
    finally:
        s.dispose()
    kinds = _kinds(pos)
    n = kinds.size
    cc, co, hd = np.zeros(n, int), np.zeros(n, int), np.zeros(n, int)
    for i, j in edges:
        ki, kj = kinds[i], kinds[j]
        if ki == carbon and kj == carbon:
            cc[i] += 1
            cc[j] += 1
        elif ki == carbon and kj == head:
            co[i] += 1
            hd[j] += 1
        elif kj == carbon and ki == head:
            co[j] += 1
            hd[i] += 1
    print(f"VALENCE N={n} bonds={len(edges)} maxCC={cc[kinds == carbon].max()} maxCO={co[kinds == carbon].max()} "
          f"maxHead={hd[kinds == head].max()}", flush=True)
    assert len(edges) > 0
    assert cc[kinds == carbon].max() <= 2
    assert co[kinds == carbon].max() <= 1
    assert hd[kinds == head].max() <= soup["headPlacement"]["chainCapacity"]

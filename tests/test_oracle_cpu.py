"""Pins the CPU oracle itself (CPU only, no GPU).

Every check is independent of the oracle's own formula: brute-force sets, closed forms of the
potentials, -grad U by finite differences, exact Boltzmann enumeration, Henderson-Hasselbalch,
equipartition of the discretised Langevin thermostat, and the reference's own invariants.
"""
import math

import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup.electrostatics import PcgState, esPairEnergy, hendersonAlpha, initialCharges, makeEsBasis, pcgNext
from protocell_soup.init_state import lattice_state, mulberry32_draws
from protocell_soup.rules import loadParams, loadSoup
from soup_fixtures import BROTH, amphiphile_opts, charged

SOUP = loadSoup()
P = loadParams()
KIND = {m["id"]: k for k, m in enumerate(SOUP["monomers"])}


def _state(opts, relax=False):
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0, prng=R.prng, next_sweep_at=R.next_sweep_at)
    if relax:
        assert o.relax(200, 0.1) == 0
    o.rebuild()
    o.force_list()
    return R, o


# ---------------------------------------------------------------------------------------------
# RNG streams
# ---------------------------------------------------------------------------------------------
def _pcg_np(v):
    v = np.asarray(v, dtype=np.uint64)
    s = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def test_pcg_hash_stream_and_uniform_map():
    xs = np.random.default_rng(0).integers(0, 2 ** 32, 2000, dtype=np.uint64)
    assert all(orc.pcg_soup(int(x)) == int(y) for x, y in zip(xs, _pcg_np(xs)))
    # iterated hash: the stream is s_{k+1} = pcg(s_k); its outputs are uniform on [0, 2^32)
    s, out = 12345, []
    for _ in range(200000):
        s = orc.pcg_soup(s)
        out.append(s)
    u = np.array(out, dtype=np.float64) / 2 ** 32
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
    # u32 -> f32 rounds to nearest: 2^32 - 128 and above map to exactly 1.0f
    assert np.float32(np.uint32(2 ** 32 - 128)) / np.float32(2 ** 32) == 1.0
    assert np.float32(np.uint32(2 ** 32 - 129)) / np.float32(2 ** 32) < 1.0


def test_cpu_pcg_next_matches_host_and_reference_stream_rule():
    st = 987654321
    host = PcgState(st)
    for _ in range(5000):
        v, st = orc.pcg_next(st)
        assert v == pcgNext(host) and st == host.state
    # the state is an LCG; the output is a permutation of the NEW state
    s0 = 1
    v, s1 = orc.pcg_next(s0)
    assert s1 == (s0 * 747796405 + 2891336453) % 2 ** 32


def test_mulberry32_known_sequence_and_lattice():
    # mulberry32 is a published generator (Tommy Ettinger); the host's vectorised form and the
    # oracle's scalar form are independent implementations of it
    lat_host = lattice_state((16.0, 16.0, 16.0), [600, 100, 500, 20, 0, 0], 7, 1.0)
    lat_orc = orc.lattice((16.0, 16.0, 16.0), [600, 100, 500, 20, 0, 0], 7, 1.0)
    for k in ("positions", "velocities", "bondRng", "thermoRng"):
        assert np.array_equal(lat_host[k], lat_orc[k]), k
    pos = lat_host["positions"].reshape(-1, 4)
    nx = 11  # ceil(cbrt(1220))
    sp = 16.0 / nx
    site = np.floor(pos[:, :3] / sp).astype(int)
    assert len({tuple(s) for s in site}) == pos.shape[0]  # each lattice site used once
    frac = pos[:, :3] / sp - site - 0.5
    assert np.abs(frac).max() <= 0.15 + 1e-5
    # bondRng0[i] = seed ^ imul(i+1, 2654435761) ^ 0x9e3779b9 (soup-init-state.ts:270)
    assert lat_host["bondRng"][0] == (7 ^ 2654435761 ^ 0x9E3779B9) & 0xFFFFFFFF
    d = mulberry32_draws(1, 3)
    assert np.all((d >= 0) & (d < 1))


# ---------------------------------------------------------------------------------------------
# Grid and lists
# ---------------------------------------------------------------------------------------------
def test_cell_assignment_and_stable_sort():
    R, o = _state(BROTH)
    pos = o.pos.reshape(-1, 4)
    dims = np.array(R.cfg.dims, dtype=np.uint32)
    box = np.array(R.cfg.box, dtype=np.float32)
    f = np.floor((pos[:, :3] / box) * dims.astype(np.float32))
    c = np.clip(f, 0, dims - 1).astype(np.int64)
    cell = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    assert np.array_equal(o.cell_of, cell)
    assert np.array_equal(np.sort(o.cell_idx), np.arange(R.N))
    key = o.cell_of[o.cell_idx].astype(np.int64) * R.N + o.cell_idx
    assert np.all(np.diff(key) > 0)  # sorted by (cell, particle index)
    assert o.cell_start[-1] == R.N and np.array_equal(np.diff(o.cell_start), np.bincount(cell, minlength=int(np.prod(dims))))


def _brute_sets(pos, box, rng, n):
    p = pos.reshape(-1, 4)[:, :3].astype(np.float32)
    out = []
    for i in range(n):
        d = p[i] - p
        d = d - np.rint(d / box) * box
        r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        m = r < rng
        m[i] = False
        out.append(set(np.nonzero(m)[0].tolist()))
    return out


def test_verlet_list_is_bruteforce_set_in_walk_order():
    R, o = _state(BROTH)
    box = np.array(R.cfg.box, dtype=np.float32)
    sets = _brute_sets(o.pos, box, np.float32(R.cfg.list_range), R.N)
    rows = o.verlet_rows()
    dims = np.array(R.cfg.dims, dtype=np.int64)
    cell_of = o.cell_of.astype(np.int64)
    for i in range(R.N):
        lst = rows[i, : o.vcount[i]].tolist()
        assert set(lst) == sets[i] and len(lst) == len(sets[i])
        # canonical order: cells in z-major walk order around i's cell, ascending index inside
        ci = np.array([cell_of[i] % dims[0], (cell_of[i] // dims[0]) % dims[1], cell_of[i] // (dims[0] * dims[1])])

        def walk_key(j):
            cj = np.array([cell_of[j] % dims[0], (cell_of[j] // dims[0]) % dims[1], cell_of[j] // (dims[0] * dims[1])])
            off = (cj - ci + 2) % dims
            return ((off[2] * 5 + off[1]) * 5 + off[0], j)
        assert lst == sorted(lst, key=walk_key)


def test_dh_list_is_bruteforce_set():
    R, o = _state(charged())
    box = np.array(R.cfg.box, dtype=np.float32)
    heads = np.nonzero(o.pos.reshape(-1, 4)[:, 3] == KIND["O"])[0]
    assert np.array_equal(o.head_idx[: o.s.nh], heads)
    p = o.pos.reshape(-1, 4)[heads, :3]
    for h in range(0, len(heads), 7):
        d = p[h] - p
        d = d - np.rint(d / box) * box
        r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        m = r < np.float32(R.cfg.es_list_range)
        m[h] = False
        assert o.es_list[h * R.cfg.es_cap: h * R.cfg.es_cap + o.es_count[h]].tolist() == heads[m].tolist()


# ---------------------------------------------------------------------------------------------
# Forces
# ---------------------------------------------------------------------------------------------
def _pair_system(kind_a, kind_b, r, charges=(0.0, 0.0), box=20.0, es=None):
    opts = {"box": [box] * 3, "seed": 1, "kT": 1.0, "clay": False, "start": {k: 0 for k in KIND}}
    opts["start"][SOUP["monomers"][kind_a]["id"]] += 1
    opts["start"][SOUP["monomers"][kind_b]["id"]] += 1
    pos = np.array([5.0, 5.0, 5.0, kind_a, 5.0 + r, 5.0, 5.0, kind_b], dtype=np.float32)
    if kind_b < kind_a:
        pos = pos.reshape(2, 4)[::-1].reshape(-1)
    opts["resume"] = {"globalStep": 0, "liveBox": [box] * 3, "positions": pos, "velocities": np.zeros(8, np.float32),
                      "bondSlots": np.full(6, 0xFFFFFFFF, np.uint32), "centerLink": np.full(2, 0xFFFFFFFF, np.uint32),
                      "centerHeldSteps": np.zeros(2, np.uint32), "desorbEvents": np.zeros(2, np.uint32),
                      "bondRng": np.array([1, 2], np.uint32), "thermoRng": np.array([3, 4], np.uint32), "events": {},
                      "charges": np.array(charges, np.float32)}
    if es:
        opts["electrostatics"] = es
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    o.rebuild()
    return R, o


def test_pair_energy_closed_forms():
    wf = 2 ** (1 / 6)
    # C-C at r = b: WCA eps (2*1 - 1)^2 = 1 and the apolar-apolar cos^2 plateau -1 below rc -> total 0
    _, o = _pair_system(KIND["C"], KIND["C"], 1.0)
    assert o.energy_f64() == pytest.approx(0.0, abs=1e-9)
    _, o = _pair_system(KIND["C"], KIND["C"], wf * 1.0 + 1.6 + 0.01)
    assert o.energy_f64() == 0.0
    # halfway up the cos^2 ramp: -cos^2(pi/4) = -1/2
    _, o = _pair_system(KIND["C"], KIND["C"], wf + 0.8)
    assert o.energy_f64() == pytest.approx(-0.5, abs=2e-7)
    # water-head (solvent-polar) depth 4/3.5 on the plateau just beyond the WCA cut
    rc = wf * 1.0
    _, o = _pair_system(KIND["O"], KIND["W"], rc - 1e-3 + 0.0)
    b_ow = 0.975
    s6 = (b_ow / (rc - 1e-3)) ** 6
    wca = (2 * s6 - 1) ** 2 if rc - 1e-3 < wf * b_ow else 0.0
    assert o.energy_f64() == pytest.approx(wca - 4 / 3.5, rel=1e-5)
    # two charged heads under DH at 0.1 M: the shifted-force energy of the reference
    es = {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.1}
    R, o = _pair_system(KIND["O"], KIND["O"], 3.5, charges=(-1.0, -1.0), es=es)
    b = makeEsBasis(SOUP, dict(P, thermostat=dict(P["thermostat"], kT=1.0)), dict(es, minBoxSigma=20.0))
    expect = esPairEnergy(float(np.float32(5 + 3.5) - np.float32(5)), 1.0, b)
    assert o.energy_f64() == pytest.approx(expect, rel=1e-6)


# ---------------------------------------------------------------------------------------------
# soupPotential (soup/src/soup-potential.ts:145-301)
# ---------------------------------------------------------------------------------------------
def _basis(R, center_link=True, charges=True):
    from protocell_soup.potential import makePotentialBasis

    return makePotentialBasis(R.soup, R.p, R.es, centerLink=center_link, charges=charges)


def _potential(R, o, center_link=True, charges=True):
    return orc.soup_potential(_basis(R, center_link, charges), o.pos, o.bond_slots, o.center_link, o.charge, R.cfg.box)


def test_soup_potential_closed_forms():
    wf = 2 ** (1 / 6)
    # C-C at r = b: WCA 1 + plateau -1; on the ramp at rc + wc/2: -1/2; beyond rc + wc: 0
    R, o = _pair_system(KIND["C"], KIND["C"], 1.0)
    assert _potential(R, o)[0] == pytest.approx(0.0, abs=1e-12)
    R, o = _pair_system(KIND["C"], KIND["C"], wf + 0.8)
    r = float(np.float32(5.0 + wf + 0.8) - np.float32(5.0))
    assert _potential(R, o)[0] == pytest.approx(-math.cos(math.pi * (r - wf) / 3.2) ** 2, rel=1e-12)
    R, o = _pair_system(KIND["C"], KIND["C"], wf + 1.6 + 0.01)
    assert _potential(R, o)[4] == 0.0
    # head-water: WCA at b = 0.975 plus the solvent-polar depth 4/3.5 on the plateau
    R, o = _pair_system(KIND["O"], KIND["W"], 1.0)
    r = float(np.float32(6.0) - np.float32(5.0))
    s6 = (0.975 / r) ** 6
    assert _potential(R, o)[0] == pytest.approx((2 * s6 - 1) ** 2 - 4 / 3.5, rel=1e-12)
    # acid-soap: a charged and a neutral head add acidSoap * attr_v; two neutral heads have no well
    R, o = _pair_system(KIND["O"], KIND["O"], 1.5, charges=(-1.0, 0.0))
    r = float(np.float32(6.5) - np.float32(5.0))
    ramp = -math.cos(math.pi * (r - wf) / 3.2) ** 2
    assert _potential(R, o)[0] == pytest.approx(2.0 * ramp, rel=1e-12)
    assert _potential(R, o, charges=False)[0] == 0.0
    # two charged heads under DH at 0.1 M: exactly esPairEnergy of the reference's basis
    es = {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.1}
    R, o = _pair_system(KIND["O"], KIND["O"], 3.5, charges=(-1.0, -1.0), es=es)
    r = float(np.float32(8.5) - np.float32(5.0))
    assert _potential(R, o)[0] == pytest.approx(esPairEnergy(r, 1.0, R.es), rel=1e-12)


def test_soup_potential_bonded_closed_forms():
    # a bent 3-bead chain C0-C1-C2 (bonds in both slots, as the reference stores them) plus a tether
    opts = {"box": [20.0] * 3, "seed": 1, "kT": 1.0, "clay": False, "start": {k: 0 for k in KIND}}
    opts["start"]["C"] = 3
    opts["start"]["M"] = 1
    pos = np.array([5.0, 5.0, 5.0, KIND["C"], 6.0, 5.0, 5.0, KIND["C"], 6.5, 5.75, 5.0, KIND["C"],
                    5.0, 6.0, 5.5, KIND["M"]], dtype=np.float32)
    none = 0xFFFFFFFF
    bs = np.array([1, none, none, 0, 2, none, 1, none, none, none, none, none], np.uint32)
    cl = np.array([3, none, none, 0], np.uint32)
    opts["resume"] = {"globalStep": 0, "liveBox": [20.0] * 3, "positions": pos, "velocities": np.zeros(16, np.float32),
                      "bondSlots": bs, "centerLink": cl, "centerHeldSteps": np.zeros(4, np.uint32),
                      "desorbEvents": np.zeros(2, np.uint32), "bondRng": np.arange(4, dtype=np.uint32),
                      "thermoRng": np.arange(4, dtype=np.uint32), "events": {}, "charges": np.zeros(4, np.float32)}
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    k, ri, kb, r0 = P["fene"]["k"], P["fene"]["rInf"], P["bend"]["k"], P["bend"]["r0"]
    fene = lambda r: -0.5 * k * ri * ri * math.log(1 - (r / ri) ** 2)  # noqa: E731
    r01, r12, r02 = 1.0, math.hypot(0.5, 0.75), math.hypot(1.5, 0.75)
    r03 = math.hypot(1.0, 0.5)
    u = _potential(R, o)
    assert u[1] == pytest.approx(fene(r01) + fene(r12), rel=1e-12)
    assert u[2] == pytest.approx(0.5 * kb * (r02 - r0) ** 2, rel=1e-12)
    assert u[3] == pytest.approx(fene(r03), rel=1e-12)
    assert _potential(R, o, center_link=False)[3] == 0.0
    assert u[4] == pytest.approx(sum(u[:4]), rel=1e-15)


def test_soup_potential_gradient_is_the_force():
    opts = amphiphile_opts(16.0, 60, 4, {"W": 1500, "C": 200, "M": 20}, seed=3,
                           electrostatics={"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1})
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    assert o.relax(200, 0.1) == 0
    o.rebuild()
    assert (o.charge != 0).sum() > 10
    f64 = o.force_f64(use_list=False)
    b = _basis(R)
    rng = np.random.default_rng(5)
    heads = np.nonzero(o.charge != 0)[0]
    picks = list(rng.choice(heads, 3, replace=False)) + list(rng.choice(R.N, 3, replace=False))
    # exact in f32 for coordinates below 16; small enough that a stencil rarely straddles the WCA cut
    # or an attraction-shell edge, where the force is continuous but its derivative jumps
    h = 2.0 ** -12
    for i in picks:
        for a in range(3):
            x0 = o.pos[4 * i + a]
            vals = []
            for kk in (-2, -1, 1, 2):
                o.pos[4 * i + a] = x0 + np.float32(kk * h)
                vals.append(orc.soup_potential(b, o.pos, o.bond_slots, o.center_link, o.charge, R.cfg.box)[4])
            o.pos[4 * i + a] = x0
            fd = -(8 * (vals[2] - vals[1]) - (vals[3] - vals[0])) / (12 * h)
            assert fd == pytest.approx(f64[i, a], rel=1e-4, abs=1e-4 * max(1.0, np.abs(f64[i]).max()))


def _fd_force(o, i, a, h=1e-3):
    """-dU/dx by a 4-point stencil on the f32 coordinate (the step actually taken is measured)."""
    x0 = o.pos[4 * i + a]
    vals, xs = [], []
    for k in (-2, -1, 1, 2):
        o.pos[4 * i + a] = np.float32(x0 + k * h)
        xs.append(float(o.pos[4 * i + a]))
        vals.append(o.energy_f64())
    o.pos[4 * i + a] = x0
    # fit a cubic through the four points and differentiate at x0
    c = np.polyfit(np.array(xs) - float(x0), vals, 3)
    return -c[2]


def test_force_is_minus_grad_energy_charged_and_bonded():
    opts = amphiphile_opts(16.0, 60, 4, {"W": 1500, "C": 200, "M": 20}, seed=3,
                           electrostatics={"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.01})
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    assert o.relax(200, 0.1) == 0
    o.rebuild()
    # give a few catalysts a tether so the centerLink FENE term is exercised
    cat = np.nonzero(o.pos.reshape(-1, 4)[:, 3] == KIND["M"])[0]
    free_c = [i for i in np.nonzero(o.pos.reshape(-1, 4)[:, 3] == KIND["C"])[0] if o.bond_slots[3 * i] == 0xFFFFFFFF]
    for m in cat[:5]:
        d = o.pos.reshape(-1, 4)[free_c, :3] - o.pos.reshape(-1, 4)[m, :3]
        d -= 16.0 * np.rint(d / 16.0)
        j = free_c[int(np.argmin((d ** 2).sum(1)))]
        if o.center_link[j] == 0xFFFFFFFF and np.sqrt((d ** 2).sum(1).min()) < 1.4:
            o.center_link[m], o.center_link[j] = j, m
    f64 = o.force_f64(use_list=False)
    assert (o.charge != 0).sum() > 10
    rng = np.random.default_rng(1)
    heads = np.nonzero(o.charge != 0)[0]
    picks = list(rng.choice(heads, 4, replace=False)) + list(rng.choice(R.N, 4, replace=False))
    picks += [int(np.nonzero(o.center_link != 0xFFFFFFFF)[0][0])] if (o.center_link != 0xFFFFFFFF).any() else []
    for i in picks:
        for a in range(3):
            fd = _fd_force(o, int(i), a)
            assert fd == pytest.approx(f64[i, a], rel=2e-4, abs=2e-4 * max(1.0, np.abs(f64[i]).max()))


def test_list_force_equals_bruteforce_and_f32_accuracy():
    R, o = _state(charged())
    fl = o.force_f64(use_list=True)
    fb = o.force_f64(use_list=False)
    assert np.abs(fl - fb).max() < 1e-9 * max(1.0, np.abs(fb).max())
    # the reference-order f32 arithmetic (WGSL expression order, box-absolute minimum image) stays
    # far inside the reference's own grid-vs-brute tolerance of 1e-2 (tests/soup-forces.test.ts:64)
    f32 = o.force.reshape(-1, 4)[:, :3].astype(np.float64)
    assert np.abs(f32 - fb).max() < 1e-3


# ---------------------------------------------------------------------------------------------
# Integrator
# ---------------------------------------------------------------------------------------------
def test_thermostat_equipartition():
    # dilute gas: the velocity variance of v <- v(1 - g dt) + sqrt(2 g kT dt) xi is kT / (1 - g dt / 2)
    opts = {"box": [40, 40, 40], "seed": 9, "kT": 1.0, "clay": False, "start": {"C": 0, "O": 0, "H": 0, "M": 0, "W": 400}}
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    o.rebuild()
    o.force_list()
    samples = []
    for k in range(3000):
        o.one_step()
        if k >= 500 and k % 10 == 0:
            samples.append(o.vel.reshape(-1, 4)[:, :3].copy())
    v2 = np.mean(np.square(samples))
    expect = 1.0 / (1.0 - 0.5 * P["thermostat"]["gamma"] * P["integrator"]["dt"])
    assert v2 == pytest.approx(expect, rel=0.02)


def test_kick_drift_wraps_into_box():
    R, o = _state(BROTH)
    o.vel[0::4] = 50.0  # push particles across the x face
    o.kick_drift()
    p = o.pos.reshape(-1, 4)
    assert (p[:, :3] >= 0).all() and (p[:, :3] <= np.float32(16.0)).all()


# ---------------------------------------------------------------------------------------------
# Bond Monte Carlo
# ---------------------------------------------------------------------------------------------
def _check_bond_invariants(o, n):
    bs = o.bond_slots.reshape(-1, 3)
    for i in range(n):
        for j in bs[i]:
            if j != 0xFFFFFFFF:
                assert i in bs[j], (i, j)
    cl = o.center_link
    for i in range(n):
        if cl[i] != 0xFFFFFFFF:
            assert cl[cl[i]] == i
    kinds = o.pos.reshape(-1, 4)[:, 3]
    # valence: heads hold at most chainCapacity chains; carbons at most 2 chain + 1 head
    heads = kinds == KIND["O"]
    assert ((bs[heads] != 0xFFFFFFFF).sum(1) <= SOUP["headPlacement"]["chainCapacity"]).all()


def test_bond_mc_catalyst_gate_and_invariants():
    base = dict(BROTH, seed=3, start={"C": 900, "O": 300, "H": 300, "M": 60, "W": 0})
    R, o = _state(base)
    census = np.bincount(o.pos.reshape(-1, 4)[:, 3].astype(int), minlength=6)
    assert o.step(3000) == 0
    _check_bond_invariants(o, R.N)
    assert o.events[0] > 0  # cc_bond with catalyst present
    assert np.array_equal(np.bincount(o.pos.reshape(-1, 4)[:, 3].astype(int), minlength=6), census)
    no_cat = dict(BROTH, seed=3, start={"C": 900, "O": 300, "H": 300, "M": 0, "W": 0})
    R2, o2 = _state(no_cat)
    assert o2.step(3000) == 0
    assert o2.events[0] == 0  # cc_bond needs a catalyst (tests/soup-bonds.test.ts:10-52)


# ---------------------------------------------------------------------------------------------
# Constant-pH Monte Carlo (tests/soup-electrostatics.test.ts:149-242)
# ---------------------------------------------------------------------------------------------
def _two_heads(r, box, pH):
    opts = {"box": [box] * 3, "seed": 1, "kT": 1.0, "clay": False,
            "start": {"C": 0, "O": 2, "H": 0, "M": 0, "W": 0},
            "electrostatics": {"enabled": True, "pH": pH}}
    pos = np.array([10, 10, 10, KIND["O"], 10 + r, 10, 10, KIND["O"]], dtype=np.float32)
    opts["resume"] = {"globalStep": 0, "liveBox": [box] * 3, "positions": pos, "velocities": np.zeros(8, np.float32),
                      "bondSlots": np.full(6, 0xFFFFFFFF, np.uint32), "centerLink": np.full(2, 0xFFFFFFFF, np.uint32),
                      "centerHeldSteps": np.zeros(2, np.uint32), "desorbEvents": np.zeros(2, np.uint32),
                      "bondRng": np.array([1, 2], np.uint32), "thermoRng": np.array([3, 4], np.uint32), "events": {},
                      "charges": np.zeros(2, np.float32)}
    R = ps.resolve(opts)
    return R, orc.OracleState(R.cfg, R.init, np.zeros(2, np.float32))


def test_protonation_henderson_hasselbalch_at_zero_work():
    worst = 0.0
    for pH in (3.9, 4.4, 4.9, 5.4, 5.9):
        R, o = _two_heads(100.0, 200.0, pH)
        o.s.prng = 12345
        on = 0
        for _ in range(20000):
            o.protonation_sweep()
            on += int((o.charge != 0).sum())
        worst = max(worst, abs(on / 40000 - hendersonAlpha(pH, 4.9)))
    assert worst < 0.01


def test_protonation_detailed_balance_two_heads():
    r = 1.1
    R, o = _two_heads(r, 40.0, 4.9)
    o.s.prng = 987654321
    b = R.es
    q = b.chargeDeprotonated
    u_both = esPairEnergy(r, q * q, b)
    dg = b.kT * math.log(10) * (b.pKaIntrinsic - b.pH)
    w_s = b.acidSoap
    u_pair = 0.0
    if w_s.scale > 0:
        u_pair = -w_s.scale * w_s.epsilon if r < w_s.rcAttr else -w_s.scale * w_s.epsilon * math.cos(math.pi * (r - w_s.rcAttr) / (2 * w_s.wc)) ** 2
    w = [1.0, math.exp(-(dg + u_pair) / b.kT), math.exp(-(dg + u_pair) / b.kT), math.exp(-(2 * dg + u_both) / b.kT)]
    exact = np.array(w) / sum(w)
    counts = np.zeros(4)
    for _ in range(200000):
        o.protonation_sweep()
        counts[int(o.charge[0] != 0) + 2 * int(o.charge[1] != 0)] += 1
    m = counts / counts.sum()
    worst = max(abs(m[0] - exact[0]), abs(m[1] + m[2] - exact[1] - exact[2]), abs(m[3] - exact[3]))
    assert worst < 5e-3


def test_initial_charges_host_equals_oracle():
    R = ps.resolve(charged())
    ch_o, rng_o = orc.initial_charges(R.cfg, R.init["positions"], R.seed)
    ch_h, rng_h = initialCharges(R.init["positions"], R.es, R.seed)
    assert np.array_equal(ch_o, ch_h) and rng_o == rng_h.state


# ---------------------------------------------------------------------------------------------
# Guards and relax
# ---------------------------------------------------------------------------------------------
def test_nonfinite_guard_and_relax():
    R, o = _state(BROTH)
    assert o.nonfinite() == (0, 0)
    o.pos[4 * 5] = np.nan
    o.vel[4 * 7 + 2] = np.inf
    assert o.nonfinite() == (1, 1)
    R, o = _state(dict(BROTH, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000}))
    f0 = np.abs(o.force).max()
    assert o.relax(200, 0.1) == 0
    assert np.abs(o.force).max() < 0.05 * f0


# ---------------------------------------------------------------------------------------------
# forceCandidateStatsDEBUG's cell walk
# ---------------------------------------------------------------------------------------------
def _brute_pairs_in_range(cfg, pos):
    """Ordered pairs i != j inside the pair's WCA cut or inside the attraction shell of an attracting
    class pair, over all pairs, in the same f32 operations (IEEE, no contraction)."""
    f32 = np.float32
    p = pos.reshape(-1, 4)
    kinds = p[:, 3].astype(np.int64)
    box = [f32(cfg.box[a]) for a in range(3)]
    rad = np.array([cfg.radius[k] for k in range(8)], np.float32)
    cls = np.array([cfg.cls[k] for k in range(8)], np.int64)
    attr = np.array([[cfg.attr[i][j] for j in range(4)] for i in range(4)], np.float32)
    rc = f32(cfg.wca_factor) * f32(cfg.b_tt)
    rc_hi = rc + f32(cfg.wc)
    total = 0
    for i in range(p.shape[0]):
        r2 = np.zeros(p.shape[0], np.float32)
        for a in range(3):
            d = p[i, a] - p[:, a]
            m = d - np.rint(d / box[a]) * box[a]
            r2 = r2 + m * m
        r = np.sqrt(r2)
        b = (f32(cfg.sigma) * (rad[kinds[i]] + rad[kinds])) * f32(0.5)
        wca = r < f32(cfg.wca_factor) * b
        att = (attr[cls[kinds[i]], cls[kinds]] > 0) & (r >= rc) & (r <= rc_hi)
        hit = wca | att
        hit[i] = False
        total += int(hit.sum())
    return total


def test_force_candidate_stats_against_brute_force():
    # a 5-cell box with walk radius 2 walks every cell: each particle examines all N - 1 others
    R, o = _state(dict(BROTH, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 1000}))
    assert list(R.cfg.dims) == [5, 5, 5] and R.cfg.walk_radius == 2
    cand, pairs = o.force_candidate_stats()
    assert cand == R.N * (R.N - 1)
    assert pairs == _brute_pairs_in_range(R.cfg, o.pos)
    # a larger box walks 5^3 of 7^3 cells: the candidates are the particles of those cells
    R, o = _state({"box": [21, 21, 21], "seed": 3, "kT": 1.0, "clay": False, "start": {"C": 900, "O": 300, "H": 300, "M": 20, "W": 0}})
    assert list(R.cfg.dims) == [7, 7, 7]
    cand, pairs = o.force_candidate_stats()
    cell = o.cell_of.astype(np.int64)
    cx, cy, cz = cell % 7, (cell // 7) % 7, cell // 49
    exp = 0
    for i in range(R.N):
        dx = np.minimum((cx - cx[i]) % 7, (cx[i] - cx) % 7)
        dy = np.minimum((cy - cy[i]) % 7, (cy[i] - cy) % 7)
        dz = np.minimum((cz - cz[i]) % 7, (cz[i] - cz) % 7)
        exp += int(((dx <= 2) & (dy <= 2) & (dz <= 2)).sum()) - 1
    assert cand == exp
    # every pair in range lies within the walk (the cells are wider than any cut), so the brute count holds
    assert pairs == _brute_pairs_in_range(R.cfg, o.pos)

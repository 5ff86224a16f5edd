"""The step path's loud failures on the GPU, restated from the reference's own tests.

- tests/soup-nonfinite-guard.test.ts: a 12^3 water lattice in box 20 with 20 pairs pulled 1e-4 sigma
  apart diverges in one step (WCA at r = 1e-4 overflows f32); step(1) must throw the non-finite
  guard's message naming the step, the interval and both counts out of 3N, and the healthy control
  (the same lattice without the overlaps) must stay finite over 2,000 steps.
- tests/verlet-capacity.test.ts:81-144: liquid water at 0.8 in box 30 with listCapacity 64 throws the
  overflow message at the first chunk end; the derived capacity (1126) holds every row over 100
  steps; and a list the device cannot hold is refused at creation with the byte count and the limit.
- The DH-list overflow guard (soup/src/soup-grid-verlet.ts:101-114), with a pinned small capacity.
- The drift guard (soup/src/soup-grid-verlet.ts:115-137): the measured drift maximum equals the
  oracle's restatement of soup_max_drift_main over the same positions, bit for bit, and a position
  write that moves a particle past skin/2 stops the next step with the drift message.

Every list write in the kernels is bounded by the row capacity and every cell index of a non-finite
position is clamped into the grid, so the kernels exit normally in each case and only the host
guards fire.
"""
import re

import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup.native import SOUP_E_DRIFT, SOUP_E_NOMEM, SOUP_E_NONFINITE, SOUP_E_OVERFLOW, SoupError
from protocell_soup.rules import loadSoup
from soup_fixtures import charged

pytestmark = pytest.mark.gpu

WATER_LATTICE = 12
OVERLAP_PAIRS = 20
OVERLAP_SEPARATION = 1e-4


def _water_lattice_opts(bad: int):
    """tests/soup-nonfinite-guard.test.ts's build(bad): a clean water lattice at spacing 20/12, with
    `bad` pairs of it pulled OVERLAP_SEPARATION apart along x."""
    soup = loadSoup()
    water_kind = next(k for k, m in enumerate(soup["monomers"]) if m["id"] == soup["solvent"]["waterId"])
    box = [20.0, 20.0, 20.0]
    n = WATER_LATTICE ** 3
    s = box[0] / WATER_LATTICE
    pos = np.zeros((n, 4), dtype=np.float32)
    k = 0
    for iz in range(WATER_LATTICE):
        for iy in range(WATER_LATTICE):
            for ix in range(WATER_LATTICE):
                pos[k] = [(ix + 0.5) * s, (iy + 0.5) * s, (iz + 0.5) * s, water_kind]
                k += 1
    for q in range(bad):
        a, b = 2 * q, 2 * q + 1
        pos[b, 0] = np.float32(pos[a, 0] + np.float32(OVERLAP_SEPARATION))
        pos[b, 1] = pos[a, 1]
        pos[b, 2] = pos[a, 2]
    start = {m["id"]: 0 for m in soup["monomers"]}
    start[soup["monomers"][water_kind]["id"]] = n
    resume = {"globalStep": 0, "liveBox": box, "positions": pos.reshape(-1), "velocities": np.zeros(4 * n, np.float32),
              "bondSlots": np.full(3 * n, 0xFFFFFFFF, np.uint32), "centerLink": np.full(n, 0xFFFFFFFF, np.uint32),
              "centerHeldSteps": np.zeros(n, np.uint32), "desorbEvents": np.zeros(2, np.uint32),
              "bondRng": np.arange(1, n + 1, dtype=np.uint32), "thermoRng": np.arange(1000003, 1000003 + n, dtype=np.uint32),
              "events": {}}
    return n, {"box": box, "seed": 7, "kT": 1.1, "clay": False, "start": start, "resume": resume}


def test_nonfinite_state_stops_the_run_naming_step_and_counts():
    n, healthy_opts = _water_lattice_opts(0)
    healthy = ps.createSoup(healthy_opts)
    try:
        assert healthy.nonFiniteCount() == {"pos": 0, "vel": 0}
        healthy.step(2000)
        assert healthy.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        healthy.dispose()

    _, bad_opts = _water_lattice_opts(OVERLAP_PAIRS)
    bad = ps.createSoup(bad_opts)
    try:
        assert bad.nonFiniteCount() == {"pos": 0, "vel": 0}  # a legal coordinate, not an injected NaN
        with pytest.raises(SoupError) as ei:
            bad.step(1)
        msg = str(ei.value)
        assert ei.value.code == SOUP_E_NONFINITE
        assert re.search(r"non-finite state", msg) and not re.search(r"Verlet list", msg), msg
        assert re.search(r"at step=1\b", msg), msg
        assert re.search(r"step interval 0\.\.1", msg), msg
        pos_count = int(re.search(r"positions=(\d+)", msg).group(1))
        vel_count = int(re.search(r"velocities=(\d+)", msg).group(1))
        assert pos_count > 0 and vel_count > 0, msg
        assert re.search(rf"of {n * 3}\b", msg), msg
        after = bad.nonFiniteCount()
        assert after["pos"] == pos_count and after["vel"] == vel_count
        assert bad.steps == 1
    finally:
        bad.dispose()


WATER30 = {"box": [30, 30, 30], "seed": 19, "kT": 1.1, "clay": False, "start": {"C": 0, "O": 0, "H": 0, "M": 0, "W": 21600}}


def test_verlet_capacity_overflow_is_loud_and_the_derived_capacity_holds():
    s = ps.createSoup(dict(WATER30, verletOverride={"listCapacity": 64}))
    try:
        with pytest.raises(SoupError) as ei:
            s.step(1)
        assert ei.value.code == SOUP_E_OVERFLOW
        assert re.search(r"listCapacity=64", str(ei.value)) and re.search(r"insufficient", str(ei.value)), str(ei.value)
        # the overflowing rows were cut at the capacity, inside their own row
        cnt = s.verletCount()
        assert cnt.max() == 64
    finally:
        s.dispose()

    s = ps.createSoup(WATER30)
    try:
        s.step(100)
        occ = s.listOccupancyDEBUG()
        assert s.verletConfig()["listCapacity"] == 1126
        assert occ["verletMax"] < occ["verletCap"] == 1126
        assert int((s.verletCount() >= 1126).sum()) == 0
    finally:
        s.dispose()


def test_a_list_the_device_cannot_hold_is_refused_at_creation():
    # 21,600 x 4,000,000 x 4 bytes = 345.6 GB of Verlet rows: more than this GPU's HBM. The guard
    # compares before allocating anything, so nothing is attempted.
    with pytest.raises(SoupError) as ei:
        ps.createSoup(dict(WATER30, verletOverride={"listCapacity": 4_000_000}))
    assert ei.value.code == SOUP_E_NOMEM
    msg = str(ei.value)
    assert re.search(r"345600000000 bytes", msg), msg
    assert re.search(r"device memory free=\d+ of total=\d+ bytes", msg), msg


def test_dh_list_overflow_is_loud():
    # the capacity derived for 700 heads spread over 1e9 sigma^3 is the floor of 64 (makeEsBasis,
    # soup/src/electrostatics.ts:221-229); in box 20 each head has several hundred within rc_es + skin
    opts = charged(box=20.0, seed=5, pH=7.0)
    opts["electrostatics"] = dict(opts["electrostatics"], densityVolumeSigma3=1e9)
    R = ps.resolve(opts)
    assert R.cfg.es_cap == 64
    s = ps.createSoup(opts)
    try:
        with pytest.raises(SoupError) as ei:
            s.step(1)
        assert ei.value.code == SOUP_E_OVERFLOW
        msg = str(ei.value)
        assert re.search(r"long-range electrostatics list: longRangeListCapacity=64 ", msg), msg
        assert re.search(r"was insufficient at rc_es=\d+\.\d{4} \(heads found=\d+\)", msg), msg
        assert int(s.readDEBUG("ES_COUNT").max()) == 64
    finally:
        s.dispose()


def _drift_sq_of(cfg, snap, positions_seq):
    """The oracle's soup_max_drift_main over each position set in turn, from a zero maximum."""
    o = orc.OracleState(cfg, {"positions": snap, "velocities": np.zeros_like(snap), "bondSlots": np.full(3 * cfg.n, 0xFFFFFFFF),
                              "centerLink": np.full(cfg.n, 0xFFFFFFFF), "centerHeldSteps": np.zeros(cfg.n),
                              "bondRng": np.zeros(cfg.n), "thermoRng": np.zeros(cfg.n)})
    o.pos_at_rebuild[:] = snap
    o.s.max_drift_bits = 0
    out = []
    for p in positions_seq:
        o.pos[:] = p
        o.max_drift()
        out.append(int(o.s.max_drift_bits))
    return out


def test_drift_maximum_matches_the_oracle_and_the_guard_fires():
    opts = charged(box=20.0, seed=3, pH=6.0)
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        every = R.cfg.rebuild_every
        s.step(30 + (-30 - s.steps) % every)
        s.step(1)  # the rebuild step: the snapshot is the positions after its drift
        snap = s.particles()
        seq, gpu = [], []
        for _ in range(every - 1):
            s.step(1)
            seq.append(s.particles())
            gpu.append(int(s.readDEBUG("MAX_DRIFT_SQ")[0]))
        ora = _drift_sq_of(R.cfg, snap, seq)
        assert gpu == ora
        assert 0 < np.sqrt(np.float32(np.array([gpu[-1]], np.uint32).view(np.float32)[0])) < 0.5 * R.cfg.skin

        # every particle translated 0.9 sigma along x since the rebuild (no pair distance changes):
        # the next chunk end stops the run
        assert s.steps % every == 0
        s.step(1)
        p = s.particles().reshape(-1, 4).copy()
        p[:, 0] = np.mod(p[:, 0] + np.float32(0.9), np.float32(R.cfg.box[0])).astype(np.float32)
        s.writeDEBUG("POS", p.reshape(-1))
        with pytest.raises(SoupError) as ei:
            s.step(1)
        assert ei.value.code == SOUP_E_DRIFT
        m = re.search(r"Verlet list: the measured drift (\d+\.\d{4}) exceeds skin/2=0\.7500", str(ei.value))
        assert m and float(m.group(1)) > 0.89, str(ei.value)
    finally:
        s.dispose()

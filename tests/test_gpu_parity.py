"""GPU parity: the HIP kernels against the oracle from the same state (through the C-ABI).

Bit-exact: cell assignment and order, Verlet and DH lists, integer RNG streams, integrator output
given the same force, bond-MC outcomes, constant-pH decisions. Forces: per particle within 1e-5 of
the FP64 recomputation, relative to max(|F64_i|, median_j |F64_j|).
"""
import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from soup_fixtures import BROTH, amphiphile_opts, charged, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


@pytest.fixture(scope="module")
def broth():
    R = ps.resolve(BROTH)
    s = ps.createSoup(BROTH)
    yield R, s
    s.dispose()


def test_priming_lists_match_oracle(broth):
    R, s = broth
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    o.rebuild()
    assert np.array_equal(s.readDEBUG("CELL_OF"), o.cell_of)
    assert np.array_equal(s.readDEBUG("CELL_START"), o.cell_start)
    assert np.array_equal(s.readDEBUG("CELL_IDX"), o.cell_idx)
    assert np.array_equal(s.verletCount(), o.vcount)
    assert np.array_equal(s.verletList(), o.verlet_rows())


def _stencil_totals(cell_start, dims, R):
    """Beads in each cell's (2R + 1)^3 neighbourhood (periodic)."""
    cnt = np.diff(np.asarray(cell_start, dtype=np.int64)).reshape(dims[2], dims[1], dims[0])
    tot = np.zeros_like(cnt)
    for dz in range(-R, R + 1):
        for dy in range(-R, R + 1):
            for dx in range(-R, R + 1):
                tot += np.roll(cnt, (dz, dy, dx), axis=(0, 1, 2))
    return tot.ravel()


def test_staged_and_gathered_list_builds_match_oracle():
    """The rebuild stages each cell's neighbourhood in LDS unless it holds more beads than
    verletOverride.stageCapacity, and then gathers that cell's candidates from the particle arrays.
    A capacity at the median neighbourhood sends about half the cells each way, 64 sends every cell
    to the gather path; every row stays bit-exact against the oracle in both cases and equal to the
    all-staged build's, at creation and after 60 steps."""
    base = charged(box=20.0, seed=6, pH=6.0)
    R = ps.resolve(base)
    ref = ps.createSoup(base)
    try:
        dims, rad = list(R.cfg.dims), int(R.cfg.walk_radius)
        tot = _stencil_totals(ref.readDEBUG("CELL_START"), dims, rad)
        mid = int(np.median(tot))
        assert 64 < tot.min() and (tot > mid).any() and (tot <= mid).any()
        for stage in (mid, 64):
            s = ps.createSoup(dict(base, verletOverride={"stageCapacity": stage}))
            try:
                assert s._R.cfg.list_stage == stage
                for k in range(2):
                    if k:
                        for x in (ref, s):
                            x.step(60)
                            x.forces()  # rebuild from the current positions
                    o = oracle_from_system(s, R)
                    o.rebuild()
                    assert np.array_equal(s.verletCount(), o.vcount), (stage, k)
                    assert np.array_equal(s.verletList(), o.verlet_rows()), (stage, k)
                    assert np.array_equal(s.verletList(), ref.verletList()), (stage, k)
            finally:
                s.dispose()
            ref.dispose()
            ref = ps.createSoup(base)
    finally:
        ref.dispose()


def test_forces_match_fp64(broth):
    R, s = broth
    o = oracle_from_system(s, R)
    o.rebuild()
    f = s.forces().reshape(-1, 4)[:, :3]
    f64 = o.force_f64(use_list=False)
    assert force_rel_err(f, f64) < FORCE_TOL
    fb = s.forcesBruteForce().reshape(-1, 4)[:, :3]
    assert force_rel_err(fb, f64) < FORCE_TOL


def _list_split(cfg):
    """Near/far split of a stored list row: max(force reach + skin/2, largest WCA cut + skin) + 0.002."""
    wca_max = cfg.wca_factor * cfg.sigma * max(cfg.radius)
    reach = max(wca_max, cfg.wca_factor * cfg.b_tt + cfg.wc, cfg.es_split)
    return max(reach * 1.00005 + 0.5 * cfg.skin, wca_max + cfg.skin) + 0.002


def test_forces_and_bonds_between_rebuilds():
    """A stored list row keeps its near part (closer than the split at the rebuild) at the front; the
    force reads it alone while the measured drift allows and matches the FP64 recomputation at every
    drift up to the next rebuild, and a bond pass in mid-interval, which reads only the near part,
    matches the oracle's walk of the whole row in lockstep from the same state."""
    opts = dict(BROTH, seed=3, start={"C": 900, "O": 300, "H": 300, "M": 60, "W": 0})
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        every = R.cfg.rebuild_every
        s.step(25 + (-25 - s.steps) % every)
        assert s.steps % every == 0
        s.step(1)  # this step rebuilt the list: the drift starts from 0
        # the near part of each row holds the entries closer than the split at the rebuild (those
        # within 1e-4 of it may go either way)
        pos0 = s.particles().reshape(-1, 4)[:, :3].astype(np.float64)
        rows, cnt, near = s.verletList(), s.verletCount(), s.readDEBUG("VERLET_NEAR")
        split, box = _list_split(R.cfg), np.asarray(R.cfg.box, np.float64)
        for i in range(R.N):
            d = pos0[rows[i, : cnt[i]].astype(np.int64)] - pos0[i]
            d = np.sqrt(((d - box * np.rint(d / box)) ** 2).sum(1))
            assert (d < split - 1e-4).sum() <= near[i] <= (d < split + 1e-4).sum(), i
        assert 0 < near.sum() < cnt.sum()
        drifts = []
        for k in range(1, every):
            s.step(1)
            drifts.append(float(np.sqrt(s.readDEBUG("MAX_DRIFT_SQ", np.float32)[0])))
            o = oracle_from_system(s, R)
            f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
            assert force_rel_err(f, o.force_f64(use_list=False)) < FORCE_TOL, (k, drifts[-1])
        assert drifts[-1] > 0.2, drifts
        s.step((-s.steps) % every)
        o = oracle_from_system(s, R)
        o.force[:] = s.forcesNoRebuildDEBUG()  # the first kick of the lockstep uses the current force
        for _ in range(every // 2):
            o.one_step()
        s.step(every // 2)
        assert np.array_equal(s.rngState()["bond"], o.bond_rng)
        s.runPhasesDEBUG("BONDS")
        o.bond_form()
        o.bond_break()
        assert np.array_equal(s.bondSlots(), o.bond_slots)
        assert np.array_equal(s.centerLinks(), o.center_link)
        assert np.array_equal(s.rngState()["bond"], o.bond_rng)
        assert np.array_equal(s.readDEBUG("EVENTS"), o.events[: 2 * R.cfg.n_rules])
    finally:
        s.dispose()


def test_integrator_bit_exact_given_force(broth):
    R, s = broth
    s.forces()
    o = oracle_from_system(s, R)
    o.force[:] = s.forcesNoRebuildDEBUG()
    s.runPhasesDEBUG("KICK_DRIFT")
    o.kick_drift()
    assert np.array_equal(s.particles(), o.pos)
    assert np.array_equal(s.velocities(), o.vel)
    s.runPhasesDEBUG("KICK_THERMO")
    o.kick_thermo()
    assert np.array_equal(s.rngState()["thermo"], o.thermo_rng)
    assert np.array_equal(s.velocities(), o.vel)


def test_steps_keep_rng_streams(broth):
    R, s = broth
    o = oracle_from_system(s, R)
    o.force[:] = s.forcesNoRebuildDEBUG()  # the first kick of the lockstep uses the current force
    s.step(40)
    for _ in range(40):
        o.one_step()
    assert np.array_equal(s.rngState()["thermo"], o.thermo_rng)
    assert np.array_equal(s.rngState()["bond"], o.bond_rng)
    # the trajectories separate from ulp-level force differences (chaos), slowly at first
    d = np.abs(s.particles() - o.pos).reshape(-1, 4)[:, :3]
    d = np.minimum(d, 16.0 - d)
    assert np.median(d) < 1e-3


def _bond_parity(s, R):
    s.forces()
    o = oracle_from_system(s, R)
    o.rebuild()
    assert np.array_equal(s.verletList(), o.verlet_rows())
    s.runPhasesDEBUG("BONDS")
    o.bond_form()
    o.bond_break()
    assert np.array_equal(s.bondSlots(), o.bond_slots)
    assert np.array_equal(s.centerLinks(), o.center_link)
    assert np.array_equal(s.centerHeldSteps(), o.center_held)
    assert np.array_equal(s.rngState()["bond"], o.bond_rng)
    assert np.array_equal(s.readDEBUG("EVENTS"), o.events[: 2 * R.cfg.n_rules])
    assert np.array_equal(s.readDEBUG("DESORB"), o.desorb)


def test_relax_then_step_dense():
    opts = dict(BROTH, seed=5, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000})
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        o = orc.OracleState(R.cfg, R.init, R.charges0)
        # first relax iteration from the same lattice: displacement cap 0.1 along F/|F|
        s.relaxColdStart({"iterations": 1, "maxDisplacementSigma": 0.1})
        assert o.relax(1, 0.1) == 0
        d = np.abs(s.particles() - o.pos).reshape(-1, 4)[:, :3]
        assert np.minimum(d, 16.0 - d).max() < 1e-4
        s.dispose()
        s = ps.createSoup(opts)
        res = s.relaxColdStart()
        assert res["nonFiniteAfter"] == 0 and res["maxForceAfter"] < 100.0 < res["maxForceBefore"]
        s.step(500)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()


def test_bond_mc_bit_exact():
    opts = dict(BROTH, seed=3, start={"C": 900, "O": 300, "H": 300, "M": 60, "W": 0})
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        for _ in range(6):
            _bond_parity(s, R)
            s.step(200)
        ev = s.events()
        assert ev["cc_bond"] > 0
    finally:
        s.dispose()


def test_charged_lists_far_force_and_sweep():
    opts = charged()
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        s.step(100)
        s.forces()
        o = oracle_from_system(s, R)
        o.rebuild()
        assert np.array_equal(s.readDEBUG("ES_COUNT"), o.es_count[: o.s.nh])
        es = s.readDEBUG("ES_LIST").reshape(o.s.nh, -1)
        assert np.array_equal(es, o.es_list.reshape(-1, R.cfg.es_cap)[: o.s.nh])
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        assert force_rel_err(f, o.force_f64(use_list=False)) < FORCE_TOL
        for _ in range(3):
            o = oracle_from_system(s, R)
            r_gpu = s.protonationSweepDEBUG()
            r_ref = o.protonation_sweep()
            assert np.array_equal(s.charges(), o.charge)
            assert s.protonationRngState() == o.s.prng
            assert r_gpu["accepted"] == r_ref["accepted"] and r_gpu["deprotonated"] == r_ref["deprotonated"]
            s.step(50)
    finally:
        s.dispose()


def test_dh_lists_over_the_head_grid_match_oracle():
    """A box of 42 sigma gives the DH-list build a head grid of 5 cells per axis (edge 8.4 >= (16.7 +
    0.02) / 2), so the rows come from the 125-cell walk and the ordered bitmap instead of the all-pairs
    scan: counts and rows (ascending head order) bit-exact against the oracle's all-pairs build, after
    a rebuild at the creation state and after 60 steps of motion, and one constant-pH sweep on them."""
    opts = charged(box=42.0, seed=3, pH=6.0, start={"C": 2000, "O": 2400, "H": 600, "M": 60, "W": 5000})
    R = ps.resolve(opts)
    box = float(R.cfg.box[0])
    assert int(np.floor(box / (0.5 * (float(R.cfg.es_list_range) + 0.02)))) == 5
    s = ps.createSoup(opts)
    try:
        for steps in (0, 60):
            s.step(steps)
            s.forces()
            o = oracle_from_system(s, R)
            o.rebuild()
            nh = o.s.nh
            assert nh == 2400
            counts = s.readDEBUG("ES_COUNT")
            assert np.array_equal(counts, o.es_count[:nh])
            assert counts.mean() > 100
            assert np.array_equal(s.readDEBUG("ES_LIST").reshape(nh, -1), o.es_list.reshape(-1, R.cfg.es_cap)[:nh])
        r_gpu = s.protonationSweepDEBUG()
        r_ref = o.protonation_sweep()
        assert np.array_equal(s.charges(), o.charge) and s.protonationRngState() == o.s.prng
        assert r_gpu["accepted"] == r_ref["accepted"]
    finally:
        s.dispose()


def test_amphiphile_soup_runs():
    opts = amphiphile_opts(16.0, 200, 4, {"W": 2277}, seed=1)
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        inv0 = s.invariants()
        s.step(1000)
        inv1 = s.invariants()
        assert inv0["monomers"] == inv1["monomers"]
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()


def test_potential_matches_oracle_and_force():
    """soupPotential on the GPU (soup_potential) against the oracle's FP64 restatement of
    soup/src/soup-potential.ts on the same configuration, and its finite-difference gradient against
    the GPU force, as tests/soup-area-move.test.ts checks the reference's."""
    opts = amphiphile_opts(16.0, 60, 4, {"W": 1500, "C": 200, "M": 20}, seed=3,
                           electrostatics={"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1})
    R = ps.resolve(opts)
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(100)
        pos = s.particles().reshape(-1, 4)
        cl = s.centerLinks().copy()
        kinds = pos[:, 3].astype(int)
        cat = np.nonzero(kinds == R.soup["monomers"].index(next(m for m in R.soup["monomers"] if m["id"] == "M")))[0]
        carbon = R.soup["monomers"].index(next(m for m in R.soup["monomers"] if m["id"] == "C"))
        free_c = np.nonzero((kinds == carbon) & (cl == 0xFFFFFFFF))[0]
        linked = 0
        for m in cat:
            if cl[m] != 0xFFFFFFFF:
                continue
            d = pos[free_c, :3] - pos[m, :3]
            d -= 16.0 * np.rint(d / 16.0)
            r = np.sqrt((d ** 2).sum(1))
            j = int(free_c[int(np.argmin(r))])
            if r.min() < 1.3 and cl[j] == 0xFFFFFFFF:
                cl[m], cl[j] = j, m
                linked += 1
        s.writeDEBUG("CENTER_LINK", cl)
        assert linked > 0
        u = s.potential()
        ref = orc.soup_potential(s.potentialBasis(), s.particles(), s.bondSlots(), s.centerLinks(), s.charges(), R.cfg.box)
        for k, key in enumerate(("nonbonded", "fene", "bend", "tether", "total")):
            assert u[key] == pytest.approx(ref[k], rel=1e-10, abs=1e-10), key
        assert u["fene"] > 0 and u["tether"] > 0 and u["bend"] > 0
        # -dU/dx against the GPU force of the same positions
        f = s.forces().reshape(-1, 4)
        p0 = s.particles()
        rng = np.random.default_rng(7)
        heads = np.nonzero(s.charges() != 0)[0]
        picks = list(rng.choice(heads, 2, replace=False)) + list(rng.choice(R.N, 3, replace=False)) + [int(np.nonzero(cl != 0xFFFFFFFF)[0][0])]
        h = 2.0 ** -9
        for i in picks:
            for a in range(3):
                vals = []
                for kk in (-2, -1, 1, 2):
                    p = p0.copy()
                    p[4 * i + a] += np.float32(kk * h)
                    s.writeDEBUG("POS", p)
                    vals.append(s.potential()["total"])
                fd = -(8 * (vals[2] - vals[1]) - (vals[3] - vals[0])) / (12 * h)
                assert fd == pytest.approx(float(f[i, a]), rel=1e-3, abs=1e-3 * max(1.0, float(np.abs(f[i, :3]).max())))
        s.writeDEBUG("POS", p0)
    finally:
        s.dispose()

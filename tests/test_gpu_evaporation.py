"""Solvent evaporation and rehydration on the GPU (soup/src/soup-evaporate.ts; the reference's own
tests/soup-evaporation.test.ts:142-452 restated): removal keeps every surviving particle's state bit
for bit and the force and lists of the smaller system match the oracle's; a checkpoint of the
partially evaporated state resumes onto the same trajectory; the returning beads carry the oracle's
positions-sampler stream, velocities and RNG seeds; the minimiser stays inside its analytical bound;
and one full evaporating wet -> dry -> wet cycle of data/soup.json's schedule realises the dry
density for what is left and returns the box and the census exactly."""
import math

import numpy as np
import pytest

import protocell_soup as ps
from oracle import evaporate as oev
from protocell_soup.checkpoint import decodeCheckpointResume, encodeCheckpoint
from protocell_soup.evaporate import _state
from protocell_soup.rules import loadSoup
from soup_fixtures import amphiphile_opts, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

BOX24 = [24.0, 24.0, 24.0]
START24 = {"C": 952, "O": 3808, "H": 952, "M": 63, "W": 11059}
OPTS = {"box": BOX24, "seed": 19, "kT": 1.1, "start": START24, "dryWetCycle": True, "evaporateSolvent": True, "clay": False}


def _bond_set(s):
    return set(map(tuple, s.bonds().reshape(-1, 2).tolist()))


def _valence(s):
    """Carbon-carbon <= 2, carbon-other <= 1, degree <= 3, no bond into removed territory."""
    monomers = loadSoup()["monomers"]
    pos, bonds, n = s.particles().reshape(-1, 4), s.bonds().reshape(-1, 2), s.N
    cc, co, deg = np.zeros(n, int), np.zeros(n, int), np.zeros(n, int)
    out = 0
    for i, j in bonds.tolist():
        if i >= n or j >= n:
            out += 1
            continue
        deg[i] += 1
        deg[j] += 1
        ki, kj = monomers[int(round(pos[i, 3]))]["kind"], monomers[int(round(pos[j, 3]))]["kind"]
        if ki == "carbon" and kj == "carbon":
            cc[i] += 1
            cc[j] += 1
        else:
            co[i] += ki == "carbon"
            co[j] += kj == "carbon"
    return int(((cc > 2) | (co > 1) | (deg > 3)).sum()), out


def _check_against_oracle(s):
    """The lists and the force of the live system against the oracle's rebuild of the same state.
    The force bar is 1e-5 relative to FP64, or twice the error of the oracle's own sequential f32
    evaluation where that is larger: right after an insertion the minimised liquid still holds
    contacts deep in the WCA core, whose large pair forces cancel in f32 to per-particle totals of
    order one."""
    o = oracle_from_system(s, s._R)
    o.rebuild()
    assert np.array_equal(s.verletCount(), o.vcount)
    assert np.array_equal(s.verletList(), o.verlet_rows())
    f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
    f64 = o.force_f64(use_list=False)
    o.force_list()
    floor = force_rel_err(o.force.reshape(-1, 4)[:, :3], f64)
    err = force_rel_err(f, f64)
    print(f"force error {err:.3e} (oracle f32 {floor:.3e})", flush=True)
    assert err < max(1e-5, 2.0 * floor)


def test_isolated_removal_checkpoint_and_insertion():
    s = ps.createSoup(OPTS)
    try:
        plan = s._R.evap
        s.relaxColdStart()
        s.step(3000)  # some chemistry first, so the bond graph under test is not empty
        wet = _state(s)
        inv_wet, bonds_wet, ev_wet = s.invariants(), _bond_set(s), s.events()
        assert len(bonds_wet) > 0
        half = plan.wetSolventCount // 2
        print("removal to", half, "solvent beads", flush=True)
        s.evaporateDEBUG(half)
        n_dry = plan.solventBlockStart + half
        assert s.N == n_dry and s.particles().size == 4 * n_dry
        dry = _state(s)
        per = {"positions": 4, "velocities": 4, "bondSlots": 3}
        for k, v in dry.items():
            assert np.array_equal(v, wet[k][: per.get(k, 1) * n_dry]), f"{k} changed on removal"
        inv = s.invariants()
        for m, c in START24.items():
            assert inv["monomers"][m] == (half if m == "W" else inv_wet["monomers"][m])
        assert _bond_set(s) == bonds_wet and s.events() == ev_wet
        assert _valence(s) == (0, 0)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        _check_against_oracle(s)

        # checkpoint of the partially evaporated state: the resumed system continues the trajectory
        cfg = {k: v for k, v in OPTS.items() if k != "clay"}
        file = encodeCheckpoint(s, cfg)
        assert file["N"] == n_dry and file["activeCounts"]["W"] == half
        r = ps.createSoup(dict(cfg, resume=decodeCheckpointResume(file), clay=False))
        try:
            assert r.N == n_dry and r.steps == s.steps and r.box == BOX24
            assert r.invariants()["monomers"] == inv["monomers"] and _bond_set(r) == bonds_wet
            s_pos = s.particles()
            assert np.array_equal(r.particles(), s_pos)
            print("resumed and original: 150 steps each", flush=True)
            r.step(150)
            s.step(150)
            sr = _state(r)
            for k, v in _state(s).items():
                assert np.array_equal(sr[k], v), f"resumed trajectory differs in {k}"
        finally:
            r.dispose()

        # insertion back to the whole pool, at the same box
        pre = _state(s)
        gs, start = s.steps, s.N
        ev_pre, bonds_pre = s.events(), _bond_set(s)
        count = plan.wetSolventCount - half
        print("insertion of", count, "beads", flush=True)
        rep = s.rehydrateDEBUG(plan.wetSolventCount)
        assert s.N == s._R.capacity_n and rep["inserted"] == count
        exp = oev.returning_beads(19, gs, pre["positions"], start, BOX24, count,
                                  plan.minSeparationSigma * s._R.p["sigma"], 1.1, plan.solventKind)
        assert rep["shortOfFloor"] == exp["shortOfFloor"] and rep["minAchieved"] == exp["minAchieved"]
        back = _state(s)
        # the minimiser moves positions, never velocities or RNG streams: those are as written
        v_new = back["velocities"].reshape(-1, 4)[start:]
        v_exp = exp["velocities"].reshape(-1, 4)
        assert np.count_nonzero(v_new != v_exp) <= 3 and np.abs(v_new - v_exp).max() <= np.spacing(np.float32(8.0))
        assert np.array_equal(back["velocities"][: 4 * start], pre["velocities"])
        assert np.array_equal(back["bondRng"][start:], exp["bondRng"])
        assert np.array_equal(back["thermoRng"][start:], exp["thermoRng"])
        assert (back["bondSlots"][3 * start:] == 0xFFFFFFFF).all() and (back["centerLink"][start:] == 0xFFFFFFFF).all()
        assert (back["centerHeldSteps"][start:] == 0).all() and (back["charges"][start:] == 0).all()
        p_new = back["positions"].reshape(-1, 4)[start:]
        assert (np.rint(p_new[:, 3]) == plan.solventKind).all()
        assert ((p_new[:, :3] >= 0) & (p_new[:, :3] < np.asarray(BOX24, np.float32))).all()
        assert _bond_set(s) == bonds_pre and s.events() == ev_pre and _valence(s) == (0, 0)
        assert s.invariants()["monomers"] == {**START24, "K": 0}
        # the perturbation of what was already there stays inside the descent's own bound
        assert rep["preexistingMaxDisplacement"] <= rep["displacementBound"] + 1e-6
        assert rep["preexistingRmsDisplacement"] < rep["preexistingMaxDisplacement"] + 1e-9
        assert math.isfinite(rep["maxForceAfter"]) and rep["maxForceAfter"] < rep["maxForceBefore"]
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        _check_against_oracle(s)
        print("1000 steps after rehydration", rep, flush=True)
        s.step(1000)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert s.invariants()["monomers"] == {**START24, "K": 0}
    finally:
        s.dispose()


def test_evaporating_cycle_returns_the_box_and_the_census():
    soup = loadSoup()
    dwc = soup["dryWetCycle"]
    s = ps.createSoup(OPTS)
    try:
        plan = s._R.evap
        s.relaxColdStart()
        wet_len = int(dwc["periodSteps"] * (1 - dwc["dryFraction"]))
        print("wet segment", wet_len, "steps and the wet -> dry transition", flush=True)
        s.stepCycled(wet_len)
        tr = s._R.last_transition
        assert tr["phase"] == "dry" and len(tr["rungs"]) == dwc["evaporationRampSteps"]
        n_dry = plan.organicCount + plan.drySolventCount
        assert s.cyclePhase == "dry" and s.N == n_dry
        assert s.box == list(plan.dryBox)
        assert n_dry / float(np.prod(s.box)) == pytest.approx(dwc["targetDryDensity"], abs=1e-9)
        inv = s.invariants()
        assert inv["monomers"]["W"] == plan.drySolventCount
        assert all(inv["monomers"][m] == c for m, c in START24.items() if m != "W")
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        _check_against_oracle(s)
        print("2000 dry steps", flush=True)
        s.stepCycled(2000)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        print("to the end of the period and 4000 wet steps", flush=True)
        s.stepCycled(dwc["periodSteps"] - s.steps + 4000)
        tr = s._R.last_transition
        assert tr["phase"] == "wet" and len(tr["rehydration"]) == 1
        assert tr["rehydration"][0]["inserted"] == plan.wetSolventCount - plan.drySolventCount
        assert s.box == BOX24 and s.N == s._R.capacity_n
        assert s.invariants()["monomers"] == {**START24, "K": 0} and s.invariants()["charge"] == 0
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert (s.cyclePhase, s.cycleIndex) == ("wet", 2)
    finally:
        s.dispose()


def test_charged_heads_survive_evaporation():
    # 24 sigma: the dry box of what is left (5,004 beads at 1.34) still holds the 5 cells the grid needs
    o = amphiphile_opts(24.0, 1000, 4, {"W": 6000}, seed=5,
                        electrostatics={"enabled": True, "pH": 5.5, "ionicStrengthMolar": 0.1, "sweepEverySteps": 100,
                                        "densityVolumeSigma3": 24.0 ** 3 / 4})
    o["dryWetCycle"], o["evaporateSolvent"] = True, True
    s = ps.createSoup(o)
    try:
        plan = s._R.evap
        s.relaxColdStart()
        s.step(250)
        q0 = s.charges()[: plan.solventBlockStart].copy()
        heads = s.readDEBUG("HEAD_IDX").copy()
        print("charged evaporation to 900 solvent beads", flush=True)
        s.evaporateDEBUG(900)
        assert np.array_equal(s.charges()[: plan.solventBlockStart], q0)
        assert np.array_equal(s.readDEBUG("HEAD_IDX"), heads)
        _check_against_oracle(s)
        sweeps = s.electrostatics()["sweeps"]
        s.step(300)  # a constant-pH sweep (due at the end of the call) at the smaller count
        assert s.electrostatics()["sweeps"] == sweeps + 1 and s.nonFiniteCount() == {"pos": 0, "vel": 0}
        rep = s.rehydrateDEBUG(plan.wetSolventCount)
        assert rep["inserted"] == plan.wetSolventCount - 900 and s.N == s._R.capacity_n
        _check_against_oracle(s)
        s.step(300)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()

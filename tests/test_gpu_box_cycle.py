"""Dry-wet cycling and box changes on the GPU (soup/src/soup-box-scale.ts): stepCycled against the
same schedule replayed by hand, the dry phase at exactly the dry box and density, growBoxTo (no
particle moves, lists rebuilt for the new box, no force pass), and a cohesive scaleBoxTo (rigid
clusters of proximity edges) against the host-side map and the oracle's force."""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.area_move import scaleMoleculesRigid
from protocell_soup.box_scale import CycleSchedule, applyBoxScale, mi3Distance, proximityPairs
from soup_fixtures import amphiphile_opts, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu


def _opts(seed, cycle):
    o = amphiphile_opts(20.0, 100, 4, {"W": 4580}, seed=seed)
    if cycle:
        o["dryWetCycle"] = True
        o["dryWetCycles"] = 1
    return o


def _state(s):
    return s.particles(), s.velocities(), s.bondSlots(), s.rngState()["thermo"]


def test_step_cycled_matches_the_schedule_replayed_by_hand():
    a, b = ps.createSoup(_opts(4, True)), ps.createSoup(_opts(4, True))
    try:
        for s in (a, b):
            s.relaxColdStart()
            # a short schedule on the same system (data/soup.json's is 32,000-step periods): wet for
            # steps [0, 2000), dry [2000, 4000), then settled wet
            s._R.cycle_cfg = CycleSchedule(4000, 0.5, 1)
        R = a._R
        dwc = R.soup["dryWetCycle"]
        assert (a.cyclePhase, a.cycleIndex) == ("wet", 1)
        a.stepCycled(2000)
        # the transition ramps to the dry box exactly: rampSteps increments, rampRelaxSteps between
        assert a.box == list(R.dry_box) and (a.cyclePhase, a.cycleIndex) == ("dry", 1)
        assert a.steps == 2000 + (dwc["rampSteps"] - 1) * dwc["rampRelaxSteps"]
        assert R.N / float(np.prod(R.dry_box)) == pytest.approx(dwc["targetDryDensity"], rel=1e-12)
        a.stepCycled(3000)
        assert a.box == list(R.box) and (a.cyclePhase, a.cycleIndex) == ("wet", 0)
        # by hand: step to each transition, ramp, go on
        b.step(2000)
        applyBoxScale(b, b._R.dry_box, dwc["rampSteps"], dwc["rampRelaxSteps"])
        b.step(4000 - b.steps)
        applyBoxScale(b, b._R.box, dwc["rampSteps"], dwc["rampRelaxSteps"])
        b.step(a.steps - b.steps)
        assert b.steps == a.steps
        for x, y in zip(_state(a), _state(b)):
            assert np.array_equal(x, y)
        assert a.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        a.dispose()
        b.dispose()


def test_grow_box_moves_nothing_and_rebuilds_for_the_new_box():
    # water only: growing the box without moving anything opens a gap at the old periodic
    # boundary, which a bond straddling it would not survive (FENE past r_inf); the reference uses
    # growBoxTo on an aggregate away from the boundary (soup-box-scale.ts:277-317)
    s = ps.createSoup({"box": [20, 20, 20], "seed": 6, "kT": 1.0, "clay": False,
                       "start": {"C": 0, "O": 0, "H": 0, "M": 0, "W": 5000}})
    try:
        s.relaxColdStart()
        s.step(60)
        pos0, f0 = s.particles(), s.forcesNoRebuildDEBUG()
        steps0 = s.steps
        s.growBoxTo([22.0, 23.5, 21.0], 3, 0)
        assert s.box == [22.0, 23.5, 21.0] and s.steps == steps0
        assert np.array_equal(s.particles(), pos0)
        assert np.array_equal(s.forcesNoRebuildDEBUG(), f0)  # growBoxTo runs no force pass
        R = s._R
        o = oracle_from_system(s, R)
        o.rebuild()
        assert np.array_equal(s.verletCount(), o.vcount)
        assert np.array_equal(s.verletList(), o.verlet_rows())
        s.step(100)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()


def test_cohesive_scale_box_keeps_clusters_rigid():
    s = ps.createSoup(_opts(8, False))
    try:
        s.relaxColdStart()
        s.step(60)
        before, old = s.particles(), tuple(s.box)
        target = (20.6, 19.7, 20.3)
        cut = 1.2
        # the amphiphile beads (the first 500 particles) as the cohesive members: the whole liquid at
        # this cutoff percolates through the periodic box, which no rigid map can keep (the
        # reference's own finding, soup-box-scale.ts:280-300)
        members = np.arange(500, dtype=np.uint32)
        s.scaleBoxTo(target, 1, 0, cohesion={"cutoff": cut, "memberIndices": members})
        # the host-side map over bonds plus every proximity edge among the members
        eligible = np.zeros(s.N, np.uint8)
        eligible[members] = 1
        prox = proximityPairs(before, old, cut, eligible)
        edges = np.concatenate([s.bonds(), prox]).astype(np.uint32)
        expect = scaleMoleculesRigid(before, edges, old, target)
        assert np.array_equal(s.particles(), expect)
        i, j = prox[0::2].astype(np.int64), prox[1::2].astype(np.int64)
        assert prox.size > 0
        assert np.abs(mi3Distance(expect, target, i, j) - mi3Distance(before, old, i, j)).max() < 1e-3
        # the force pass after the change sees the new geometry
        R = s._R
        o = oracle_from_system(s, R)
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        assert force_rel_err(f, o.force_f64(use_list=False)) < 1e-5
    finally:
        s.dispose()

"""Box changes and the resident force (the reference's tests/soup-stale-force.test.ts:28-147 and
tests/soup-grid-resize.test.ts:19-67 restated): after a ramped scaleBoxTo with no dynamics after it,
the force buffer the next kick consumes is the force of the new geometry (here bit-identical to a
fresh rebuild and force, since the sums are FP64 rounded once) and not the pre-change force; the
mid-run minimiser moves positions only; and a box change that crosses a cell-count boundary
reallocates the grid with the list force still agreeing with the brute-force one."""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.plan import planSoupGrid
from soup_fixtures import force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

START = {"C": 300, "O": 1200, "H": 300, "M": 20, "W": 3400}


def _system():
    s = ps.createSoup({"box": [16, 16, 16], "seed": 7, "kT": 1.1, "start": START, "clay": False})
    s.relaxColdStart()
    s.step(2000)
    return s


def test_box_change_leaves_the_force_of_the_new_geometry():
    s = _system()
    try:
        pre_fresh = s.forces()
        pre_resident = s.forcesNoRebuildDEBUG()
        assert np.array_equal(pre_resident, pre_fresh)  # step() ends with a force pass
        s.scaleBoxTo([15, 15, 15], 3, 0)
        post_resident = s.forcesNoRebuildDEBUG()
        post_fresh = s.forces()
        assert s.box == [15, 15, 15]
        mx = np.abs(pre_fresh).max()
        assert np.abs(post_fresh).max() > 0
        assert np.array_equal(post_resident, post_fresh)
        assert np.abs(post_resident - pre_fresh).max() > 0.1 * mx
        o = oracle_from_system(s, s._R)
        assert force_rel_err(post_resident.reshape(-1, 4)[:, :3], o.force_f64(use_list=False)) < 1e-5
    finally:
        s.dispose()


def test_mid_run_minimisation_moves_positions_only():
    s = _system()
    try:
        inv0, bonds0, ev0, steps0, box0 = s.invariants(), s.bonds().copy(), s.events(), s.steps, s.box
        m = s.minimiseNowDEBUG(29)
        assert (s.steps, s.box) == (steps0, box0)
        inv1 = s.invariants()
        assert inv1["monomers"] == inv0["monomers"] and inv1["charge"] == inv0["charge"]
        assert np.array_equal(s.bonds(), bonds0) and s.events() == ev0
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert m["iterations"] == 29 and m["maxForceAfter"] < m["maxForceBefore"] and m["displacementBound"] > 0
    finally:
        s.dispose()


def test_grid_reallocation_across_a_cell_count_boundary():
    # a dry box whose grid has fewer cells per axis than the wet one: the first transition of the
    # file's schedule reallocates the grid (soup-grid-resize.test.ts's findBracketCrossingBox idea)
    soup_start = {"C": 400, "O": 300, "H": 300, "M": 20, "W": 8000}
    box = 22.0
    R = ps.resolve({"box": [box] * 3, "seed": 13, "kT": 1.1, "start": soup_start, "dryWetCycle": True, "clay": False})
    wet_dims, dry_dims = planSoupGrid(R.box, R.start).dims, planSoupGrid(R.dry_box, R.start).dims
    assert wet_dims != dry_dims
    s = ps.createSoup({"box": [box] * 3, "seed": 13, "kT": 1.1, "start": soup_start, "dryWetCycle": True, "clay": False})
    try:
        s.relaxColdStart()
        before = s.invariants()
        dwc = s._R.soup["dryWetCycle"]
        s.stepCycled(int(dwc["periodSteps"] * (1 - dwc["dryFraction"])) + 1)
        assert s.cyclePhase == "dry" and s.box == list(R.dry_box)
        assert list(s._R.plan.dims) == list(dry_dims)
        after = s.invariants()
        assert after["monomers"] == before["monomers"] and after["charge"] == before["charge"]
        f = s.forces().reshape(-1, 4)[:, :3]
        fb = s.forcesBruteForce().reshape(-1, 4)[:, :3]
        assert np.abs(f - fb).max() < 1e-2
        o = oracle_from_system(s, s._R)
        o.rebuild()
        assert np.array_equal(s.verletList(), o.verlet_rows())
        assert force_rel_err(f, o.force_f64(use_list=False)) < 1e-5
    finally:
        s.dispose()

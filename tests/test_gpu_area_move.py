"""The area move on the GPU (soup/src/soup-area-move.ts): soupPotential of arbitrary configurations
from their own cell grid (soup_potential_at) against the oracle's FP64 brute force, the Metropolis
chain against the same chain driven by the oracle's energies, and the box change it applies
(applyBoxScaleOnce: rigid map, new grid, rebuild, force)."""
import math

import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup.area_move import moleculeCount, proposeBox, scaleMoleculesRigid
from protocell_soup.init_state import Mulberry32
from protocell_soup.plan import planSoupGrid
from soup_fixtures import amphiphile_opts, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

CHARGED = {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1}


def _oracle_energy(s, positions, box):
    r = orc.soup_potential(s.potentialBasis(), positions, s.bondSlots(), s.centerLinks(), s.charges(), box)
    return r


@pytest.mark.parametrize("es", [None, CHARGED])
def test_potential_at_matches_oracle(es):
    opts = amphiphile_opts(16.0, 80, 4, {"W": 1800, "C": 150, "M": 10}, seed=6, electrostatics=es)
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(60)
        pos, box = s.particles(), s.box
        # the current configuration: the same energy as the list-based potential and the oracle
        at = s.potentialAt(pos, box)
        ref = _oracle_energy(s, pos, box)
        cur = s.potential()
        for k, key in enumerate(("nonbonded", "fene", "bend", "tether", "total")):
            assert at[key] == pytest.approx(ref[k], rel=1e-10, abs=1e-10), key
            assert at[key] == pytest.approx(cur[key], rel=1e-10, abs=1e-10), key
        # rigidly mapped into proposed boxes of both modes (the area move's trials)
        pairs = s.bonds()
        for u, mode in ((0.011, "lateral-fixed-volume"), (-0.012, "lateral-fixed-volume"), (0.009, "lateral-fixed-z")):
            nb = proposeBox(box, u, mode)
            q = scaleMoleculesRigid(pos, pairs, box, nb)
            at = s.potentialAt(q, nb)
            ref = _oracle_energy(s, q, nb)
            for k, key in enumerate(("nonbonded", "fene", "bend", "tether", "total")):
                assert at[key] == pytest.approx(ref[k], rel=1e-10, abs=1e-10), (u, mode, key)
        if es is not None:
            assert np.count_nonzero(s.charges()) > 0
    finally:
        s.dispose()


def test_area_move_chain_matches_oracle_and_applies_the_box():
    opts = amphiphile_opts(16.0, 120, 4, {"W": 2000}, seed=8)
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(100)
        R = s._R
        start_box = tuple(s.box)
        start_pos = s.particles()
        pairs = s.bonds()
        # the oracle's replay of makeSoupAreaMove's loop (soup-area-move.ts:228-255): same stream,
        # same proposals and maps, FP64 energies from the oracle
        rng = Mulberry32((R.seed ^ 0x5EEDA1EA) & 0xFFFFFFFF)
        am = R.soup["areaMove"]
        kT = R.p["thermostat"]["kT"]
        n_mol = moleculeCount(pairs, R.N)
        cur_pos, cur_box = start_pos, start_box
        energy = _oracle_energy(s, cur_pos, cur_box)[4]
        traj, acc = [], 0
        for _ in range(12):
            u = (float(rng.take(1)[0]) * 2 - 1) * am["logDelta"]
            pb = proposeBox(cur_box, u, am["mode"])
            if not planSoupGrid(pb, R.start).valid:
                traj.append(cur_box[0])
                continue
            pp = scaleMoleculesRigid(cur_pos, pairs, cur_box, pb)
            pe = _oracle_energy(s, pp, pb)[4]
            vr = (pb[0] * pb[1] * pb[2]) / (cur_box[0] * cur_box[1] * cur_box[2])
            if float(rng.take(1)[0]) < min(1.0, math.exp(n_mol * math.log(vr) - (pe - energy) / kT)):
                acc += 1
                cur_pos, cur_box, energy = pp, pb, pe
            traj.append(cur_box[0])
        res = s.areaMove(12)
        assert res["accepted"] == acc and res["lateralTrajectory"] == traj
        assert res["box"] == list(cur_box) and 0 < acc < 12
        assert res["energyEnd"] == pytest.approx(energy, rel=1e-10)
        # the applied box: the system now holds the rigid map of the start configuration into it
        assert tuple(s.box) == cur_box
        expect = scaleMoleculesRigid(start_pos, pairs, start_box, cur_box)
        assert np.array_equal(s.particles(), expect)
        o = oracle_from_system(s, R)
        o.rebuild()
        assert np.array_equal(s.verletCount(), o.vcount)
        assert np.array_equal(s.verletList(), o.verlet_rows())
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        assert force_rel_err(f, o.force_f64(use_list=False)) < 1e-5
        # dynamics continue in the new box
        s.step(200)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        p = s.particles().reshape(-1, 4)
        assert np.all(p[:, :3] >= 0) and np.all(p[:, :3] <= np.asarray(cur_box, np.float32))
    finally:
        s.dispose()

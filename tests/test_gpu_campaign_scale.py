"""Parity at the campaign's full size: the 483,268-bead soup (76 sigma, pH 7.0, 0.01 M long-range
Debye-Hueckel) after the cold-start relax and 37 steps, against the oracle on the same state:
the cell assignment, every Verlet count and row (set and order, 2.2 GB of list), the head lists of
the long-range term, the list force within 1e-5 of the FP64 recomputation over the oracle's own
lists, the decisions of one constant-pH sweep over all 18,659 heads, and one bond Monte Carlo
phase (claims, events, desorption and RNG streams)."""
import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from soup_fixtures import force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

CAMPAIGN = {"box": [76, 76, 76], "seed": 19, "kT": 1.1, "clay": False,
            "start": {"C": 55988, "O": 18659, "H": 55988, "M": 1452, "W": 351181},
            "electrostatics": {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01}}


def test_campaign_scale_parity():
    orc.set_threads(16)
    s = ps.createSoup(CAMPAIGN)
    try:
        R = s._R
        assert s.N == 483268
        s.relaxColdStart()
        s.step(37)  # between rebuilds, past a bond Monte Carlo step
        s.forces()  # lists and force of this exact state
        o = oracle_from_system(s, R)
        o.rebuild()
        print("oracle rebuilt", flush=True)
        assert np.array_equal(s.readDEBUG("CELL_OF"), o.cell_of)
        assert np.array_equal(s.verletCount(), o.vcount)
        assert np.array_equal(s.verletList(), o.verlet_rows())
        nh = o.s.nh
        assert nh == 18659
        assert np.array_equal(s.readDEBUG("ES_COUNT"), o.es_count[:nh])
        assert np.array_equal(s.readDEBUG("ES_LIST").reshape(nh, -1), o.es_list.reshape(-1, R.cfg.es_cap)[:nh])
        print("lists equal", flush=True)
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        err = force_rel_err(f, o.force_f64(use_list=True))
        print(f"force error {err:.3e}", flush=True)
        assert err < 1e-5
        # one constant-pH sweep from the same charges and PCG state
        r_gpu = s.protonationSweepDEBUG()
        r_ref = o.protonation_sweep()
        assert np.array_equal(s.charges(), o.charge)
        assert s.protonationRngState() == o.s.prng
        assert (r_gpu["accepted"], r_gpu["deprotonated"]) == (r_ref["accepted"], r_ref["deprotonated"])
        # one bond Monte Carlo phase on this state: every claim, event and RNG draw
        s.runPhasesDEBUG("BONDS")
        o.bond_form()
        o.bond_break()
        assert np.array_equal(s.bondSlots(), o.bond_slots)
        assert np.array_equal(s.centerLinks(), o.center_link)
        assert np.array_equal(s.centerHeldSteps(), o.center_held)
        assert np.array_equal(s.rngState()["bond"], o.bond_rng)
        assert np.array_equal(s.readDEBUG("EVENTS"), o.events[: 2 * R.cfg.n_rules])
        assert np.array_equal(s.readDEBUG("DESORB"), o.desorb)
        print("bond phase equal, events", s.events(), flush=True)
    finally:
        s.dispose()

"""Dry-wet cycling at liquid water against a matched control (the reference's
tests/soup-drywet-cycling.test.ts:85-244 restated): one full cycle of data/soup.json's schedule
returns the box exactly and stays finite, the ramp takes (rampSteps - 1) * rampRelaxSteps steps,
and the dry phase suppresses mobility (mean-square displacement below half the control's over the
same window) and chemistry (bond events per step below the control's, and fewer over the cycle)."""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.rules import loadSoup

pytestmark = pytest.mark.gpu

BOX = [20.0, 20.0, 20.0]
START = {"C": 551, "O": 2204, "H": 551, "M": 37, "W": 6400}
WINDOW = 2000


def _msd(p0, p1, box):
    d = p1.reshape(-1, 4)[:, :3].astype(np.float64) - p0.reshape(-1, 4)[:, :3].astype(np.float64)
    L = np.asarray(box, np.float64)
    r = np.floor(d / L)
    m = d - (r + ((d / L) - r >= 0.5)) * L
    return float(((m * m).sum(axis=1)).mean())


def _events(s):
    e = s.events()
    return e["cc_bond"] + e["co_bond"]


def test_dry_phase_suppresses_mobility_and_chemistry():
    dwc = loadSoup()["dryWetCycle"]
    wet_len = int(dwc["periodSteps"] * (1 - dwc["dryFraction"]))
    base = {"box": BOX, "seed": 19, "kT": 1.1, "start": START, "clay": False}
    cyc, ctl = ps.createSoup(dict(base, dryWetCycle=True)), ps.createSoup(base)
    try:
        for s in (cyc, ctl):
            s.relaxColdStart()
            s.stepCycled(wet_len - WINDOW - 200)
        # identical up to here: the cycle has not started yet
        assert np.array_equal(cyc.particles(), ctl.particles())
        wp0, we0 = cyc.particles(), _events(cyc)
        cyc.stepCycled(WINDOW)
        msd_wet = _msd(wp0, cyc.particles(), cyc.box)
        wet_rate = (_events(cyc) - we0) / WINDOW
        before = cyc.steps
        cyc.stepCycled(200)  # crosses the wet -> dry transition
        ramp = cyc.steps - before - 200
        assert ramp == (dwc["rampSteps"] - 1) * dwc["rampRelaxSteps"]
        n = sum(START.values())
        assert n / float(np.prod(cyc.box)) == pytest.approx(dwc["targetDryDensity"], abs=1e-5)
        dp0, de0 = cyc.particles(), _events(cyc)
        cyc.stepCycled(WINDOW)
        msd_dry = _msd(dp0, cyc.particles(), cyc.box)
        dry_rate = (_events(cyc) - de0) / WINDOW
        ctl.stepCycled(cyc.steps - WINDOW - ctl.steps)
        cp0, ce0 = ctl.particles(), _events(ctl)
        ctl.stepCycled(WINDOW)
        msd_ctl = _msd(cp0, ctl.particles(), ctl.box)
        ctl_rate = (_events(ctl) - ce0) / WINDOW
        cyc.stepCycled(dwc["periodSteps"] - cyc.steps + ramp)
        assert cyc.box == BOX and cyc.cyclePhase == "wet"
        ctl.stepCycled(cyc.steps - ctl.steps)
        ev_cyc, ev_ctl = _events(cyc), _events(ctl)
        print(f"msd wet={msd_wet:.4f} dry={msd_dry:.4f} control={msd_ctl:.4f}; event rate wet={wet_rate:.5f} "
              f"dry={dry_rate:.5f} control={ctl_rate:.5f}; cycle total cycled={ev_cyc} control={ev_ctl}", flush=True)
        for s in (cyc, ctl):
            assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert cyc.invariants()["monomers"] == ctl.invariants()["monomers"]
        assert msd_dry < msd_ctl / 2
        assert dry_rate < ctl_rate
        assert ev_cyc < ev_ctl
    finally:
        cyc.dispose()
        ctl.dispose()

"""The displacement-triggered Verlet rebuild (verletOverride.rebuild = "drift") on the GPU.

The reference rebuilds on a fixed cadence (soup/src/soup-integrate.ts:149) and only asserts the
measured drift (soup/wgsl/verlet.wgsl:115-123, soup/src/soup-grid-verlet.ts:125-136). The drift
schedule reads that drift maximum back every driftCheckEvery steps and rebuilds the next step once it
exceeds the trigger. Its lists differ from the cadence's on the steps where the schedules differ, so
it is checked by what the lists are for: the force after every step is within 1e-5 of the FP64 force
over all pairs (no list), the drift never passes skin/2, and the rebuild count follows the drift.
Replaying its segments from captured graphs must equal stepping kernel by kernel bit for bit; the
cadence schedule's count is exactly steps / rebuildEvery.
"""
import numpy as np
import pytest

import protocell_soup as ps
from soup_fixtures import BROTH, assert_same_state, charged, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu


def _drift(opts, every=5):
    o = dict(opts)
    o["verletOverride"] = dict(o.get("verletOverride") or {}, rebuild="drift", driftCheckEvery=every)
    return o


def _drift_now(s):
    return float(np.sqrt(s.readDEBUG("MAX_DRIFT_SQ", np.float32)[0]))


def test_drift_schedule_keeps_every_force_within_tolerance():
    opts = _drift(charged(box=20.0, seed=6, pH=6.0), every=1)
    R = ps.resolve(opts)
    assert R.cfg.rebuild_mode == 1 and R.cfg.drift_check_every == 1
    trig = float(R.cfg.drift_trigger)
    assert 0.0 < trig < 0.5 * R.cfg.skin
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart({"iterations": 50})
        s.step(100)
        r0 = s.rebuildCount()
        drifts, rebuilt = [], []
        for k in range(60):
            before = s.rebuildCount()
            s.step(1)
            rebuilt.append(s.rebuildCount() - before)
            drifts.append(_drift_now(s))
            if (s.steps - 1) % R.cfg.bond_interval == 0:
                continue  # the step's force preceded its bond phase, which may have changed the bonds
            o = oracle_from_system(s, R)
            err = force_rel_err(s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3], o.force_f64(use_list=False))
            assert err < 1e-5, (k, drifts[-1], err)
        assert max(drifts) <= 0.5 * R.cfg.skin
        # a rebuild follows every check that found the drift past the trigger, and only those
        for k in range(1, 60):
            assert rebuilt[k] == (1 if drifts[k - 1] > trig else 0), (k, drifts[k - 1], rebuilt[k])
        assert sum(rebuilt) == s.rebuildCount() - r0 > 0
        print(f"DRIFT trigger={trig:.4f} rebuilds={sum(rebuilt)} in 60 steps, max drift {max(drifts):.4f}", flush=True)
    finally:
        s.dispose()


def test_drift_segments_replayed_from_graphs_equal_direct_steps():
    opts = _drift(dict(BROTH, seed=4, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000}), every=5)
    a, b = ps.createSoup(opts), ps.createSoup(opts)
    try:
        b.useGraphs(False)
        for s in (a, b):
            s.relaxColdStart({"iterations": 50})
        for n in (37, 1000, 263, 5, 1):
            a.step(n)
            b.step(n)
            assert_same_state(a, b, f"after step({n})")
            assert a.rebuildCount() == b.rebuildCount()
        assert a.steps == 1306
        rate = a.steps / max(1, a.rebuildCount())
        print(f"DRIFT graphs: {a.rebuildCount()} rebuilds in {a.steps} steps (one per {rate:.1f})", flush=True)
        assert 1.0 < rate < 100.0
    finally:
        a.dispose()
        b.dispose()


def test_long_drift_run_and_cadence_count():
    opts = dict(BROTH, seed=8, kT=1.1, start={"C": 600, "O": 200, "H": 500, "M": 20, "W": 2500})
    d = ps.createSoup(_drift(opts))
    c = ps.createSoup(opts)
    try:
        for s in (d, c):
            s.relaxColdStart({"iterations": 50})
            s.step(3000)  # three chunks: every chunk-end guard (drift <= skin/2 among them) passes
            assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert c.rebuildCount() == 3000 // c._R.cfg.rebuild_every
        print(f"DRIFT 3000 steps: drift schedule {d.rebuildCount()} rebuilds, cadence {c.rebuildCount()}", flush=True)
        assert 0 < d.rebuildCount()
    finally:
        d.dispose()
        c.dispose()


def test_drift_schedule_options_are_checked():
    with pytest.raises(ValueError, match="cadence"):
        ps.resolve(dict(BROTH, verletOverride={"rebuild": "sometimes"}))
    with pytest.raises(ValueError, match="driftTrigger"):
        ps.resolve(dict(BROTH, verletOverride={"rebuild": "drift", "driftTrigger": 0.9}))

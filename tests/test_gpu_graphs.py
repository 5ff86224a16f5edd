"""Step replay from captured HIP graphs, and the batched replica step (soup_step_batch), against
plain kernel-by-kernel stepping of the same system: the same state bit for bit -- positions,
velocities, forces, bond slots, centre links, RNG streams, event counters and charges -- across
graph periods, partial periods, rebuilds that flip the double-buffered state, readbacks and
constant-pH sweeps."""
import numpy as np
import pytest

import protocell_soup as ps
from soup_fixtures import BROTH, assert_same_state, charged

pytestmark = pytest.mark.gpu

DENSE = dict(BROTH, seed=5, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000})


def _charged(box, seed, pH, ionic):
    o = charged(box=box, seed=seed, pH=pH, ionic=ionic)
    o["electrostatics"]["sweepEverySteps"] = 100
    # DH rows sized for heads gathered into a quarter of the box: these soups aggregate, and the
    # capacity from a uniform head density can overflow on some trajectories
    o["electrostatics"]["densityVolumeSigma3"] = box ** 3 / 4.0
    return o


_assert_same = assert_same_state


def _pair(opts):
    a, b = ps.createSoup(opts), ps.createSoup(opts)
    b.useGraphs(False)
    for s in (a, b):
        s.relaxColdStart({"iterations": 50})
    return a, b


@pytest.mark.parametrize("opts", [DENSE, _charged(20.0, 11, 5.0, 0.01)], ids=["broth", "charged"])
def test_graph_replay_matches_direct_steps(opts):
    a, b = _pair(opts)
    try:
        bonds0 = a.bondSlots().copy()
        # whole periods from step 0, a partial period, then a rebuild that swaps the state buffers
        # (the other graph), then mixed partial and whole periods across constant-pH sweeps
        for k, n in enumerate((40, 7, 53, 120, 1, 299)):
            a.step(n)
            b.step(n)
            _assert_same(a, b, f"after call {k} ({n} steps)")
            if k == 1:
                fa, fb = a.forces(), b.forces()
                assert np.array_equal(fa, fb)
        assert a.steps == 520
        if "electrostatics" not in opts:
            assert not np.array_equal(a.bondSlots(), bonds0), "no bond changed: the bond phase was not exercised"
        else:
            assert a.electrostatics()["sweeps"] == 3  # at the ends of the calls reaching steps 100, 220 and 520
    finally:
        a.dispose()
        b.dispose()


def test_step_replicas_matches_each_system_stepped_alone():
    # uneven replicas: sizes, boxes, pH, ionic strength and seeds differ; one steps without graphs
    optss = [DENSE, _charged(20.0, 3, 5.0, 0.01), _charged(17.0, 4, 7.5, 0.1), dict(BROTH, seed=9),
             _charged(24.0, 8, 6.0, 0.001)]
    batch = [ps.createSoup(o) for o in optss]
    solo = [ps.createSoup(o) for o in optss]
    try:
        for s in batch + solo:
            s.relaxColdStart({"iterations": 50})
        batch[3].useGraphs(False)
        solo[3].useGraphs(False)
        ps.stepReplicas(batch, 230)
        for s in solo:
            s.step(230)
        for r, (x, y) in enumerate(zip(batch, solo)):
            _assert_same(x, y, f"replica {r} after 230 steps")
        # replica 1 rebuilds on its own (odd buffer assignment), then all step again, past a chunk
        batch[1].forces()
        solo[1].forces()
        ps.stepReplicas(batch, 1013)
        for s in solo:
            s.step(1013)
        for r, (x, y) in enumerate(zip(batch, solo)):
            _assert_same(x, y, f"replica {r} after 1243 steps")
            assert x.steps == 1243
    finally:
        for s in batch + solo:
            s.dispose()


def test_step_replicas_rejects_a_system_listed_twice():
    s = ps.createSoup(dict(BROTH, seed=2))
    try:
        with pytest.raises(RuntimeError):
            ps.stepReplicas([s, s], 10)
        assert s.steps == 0
        ps.stepReplicas([s], 10)
        assert s.steps == 10
    finally:
        s.dispose()


def test_profiled_stepping_counts_every_launch():
    # profiled steps run kernel by kernel (phase brackets are host-recorded events); switching profiling
    # on, changing its mask and switching it off again leaves the state bit-identical to direct stepping
    a, b = _pair(DENSE)
    try:
        for s in (a, b):
            s.profile(True, phases=["force", "bonds"])
        for n in (40, 7, 53, 120):
            a.step(n)
            b.step(n)
        _assert_same(a, b, "profiled")
        pa, pb = a.profileRead(), b.profileRead()
        for ph in ("force", "bonds"):
            assert pa[ph]["count"] == pb[ph]["count"] > 0 and pa[ph]["ms"] > 0
        assert pa["force"]["count"] == 220
        for s in (a, b):
            s.profile(True, phases=["force"])
            s.step(60)
        _assert_same(a, b, "new mask")
        assert a.profileRead()["force"]["count"] == 60 and a.profileRead()["bonds"]["count"] == 0
        for s in (a, b):
            s.profile(False)
            s.step(40)
        _assert_same(a, b, "profiling off")
    finally:
        a.dispose()
        b.dispose()

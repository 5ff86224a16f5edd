"""The SoupSystem's diagnostic hooks and runUntil on the GPU (soup/src/soup-readback.ts:191-263,
soup/src/sim.ts:452-528): forceCandidateStatsDEBUG equals the oracle's restatement of the cell walk
and leaves the trajectory untouched; stepPhasesDEBUG times its five groups and leaves consistent
lists and forces; minimiseNowDEBUG moves particles inside the descent's bound without moving the step
counter; runUntil's trace equals detectStage taken by hand on an identical twin."""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.analysis import detectStage
from protocell_soup.evaporate import _state
from soup_fixtures import BROTH, amphiphile_opts, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

DENSE = dict(BROTH, seed=5, start={"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000})


def _same(a, b, what):
    sa, sb = _state(a), _state(b)
    sa["force"], sb["force"] = a.forcesNoRebuildDEBUG(), b.forcesNoRebuildDEBUG()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), f"{what}: {k} differs"


def test_force_candidate_stats_match_the_oracle_and_touch_nothing():
    a, b = ps.createSoup(DENSE), ps.createSoup(DENSE)
    try:
        for s in (a, b):
            s.relaxColdStart({"iterations": 50})
            s.step(137)  # moved since the last rebuild: the walk sees the grid of the current positions
        st = a.forceCandidateStatsDEBUG()
        o = oracle_from_system(a, a._R)
        cand, pairs = o.force_candidate_stats()
        print("candidates", cand, "pairs", pairs, flush=True)
        assert (st["candidatesExamined"], st["pairsWithinRange"]) == (cand, pairs)
        assert st["ratio"] == cand / pairs and pairs > 0 and cand > pairs
        _same(a, b, "right after the stats")
        a.step(100)
        b.step(100)
        _same(a, b, "100 steps later")
    finally:
        a.dispose()
        b.dispose()


def test_step_phases_time_each_group_and_leave_consistent_lists():
    s = ps.createSoup(DENSE)
    try:
        s.relaxColdStart({"iterations": 50})
        s.step(40)
        ph = s.stepPhasesDEBUG(5)
        assert set(ph) == {"full", "gridBuild", "force", "bondAttempts", "integration"}
        assert all(v > 0 for v in ph.values())
        assert ph["full"] > ph["force"] and ph["full"] > ph["integration"]
        assert s.steps == 40
        o = oracle_from_system(s, s._R)
        o.rebuild()
        assert np.array_equal(s.verletList(), o.verlet_rows())
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        assert force_rel_err(f, o.force_f64(use_list=False)) < 1e-5
        s.step(200)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()


def test_minimise_now_stays_inside_its_bound():
    s = ps.createSoup(DENSE)
    try:
        s.relaxColdStart({"iterations": 50})
        s.step(200)
        x0 = s.particles().reshape(-1, 4)[:, :3].astype(np.float64)
        rep = s.minimiseNowDEBUG(30)
        assert rep["globalStep"] == 200 == s.steps and rep["iterations"] == 30
        assert rep["maxForceAfter"] <= rep["maxForceBefore"]
        x1 = s.particles().reshape(-1, 4)[:, :3].astype(np.float64)
        L = np.asarray(s.box, np.float64)
        d = x1 - x0
        d -= np.rint(d / L) * L
        assert np.sqrt((d * d).sum(axis=1)).max() <= rep["displacementBound"] + 1e-5
        s.step(100)
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()


def _amph():
    return amphiphile_opts(16.0, 120, 4, {"W": 2400}, seed=4)


def test_run_until_trace_equals_detect_stage_on_a_twin():
    a, b = ps.createSoup(_amph()), ps.createSoup(_amph())
    try:
        for s in (a, b):
            s.relaxColdStart({"iterations": 50})
        lines = []
        res = a.runUntil("vesicle", {"maxSteps": 250, "sampleEvery": 100}, log=lines.append)
        assert not res["reached"] and res["steps"] == 250 and len(res["trace"]) == 3 and len(lines) == 3
        for entry, n in zip(res["trace"], (100, 100, 50)):
            b.stepCycled(n)
            stage, ev = detectStage(b)
            assert entry["steps"] == b.steps and entry["stage"] == stage and entry["evidence"] == ev
            assert (entry["cyclePhase"], entry["cycleIndex"]) == ("none", 0)
        first = res["trace"][0]["stage"]
    finally:
        a.dispose()
        b.dispose()
    c = ps.createSoup(_amph())
    try:
        c.relaxColdStart({"iterations": 50})
        res = c.runUntil(first, {"maxSteps": 250, "sampleEvery": 100})
        assert res["reached"] and res["steps"] == 100 and len(res["trace"]) == 1
    finally:
        c.dispose()

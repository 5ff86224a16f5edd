"""Dry-wet cycling and box changes, host side (no GPU): the schedule and dry box against the
reference's own known answers (tests/soup-boxcycle.test.ts:150-179), the creation-time cycle
config and its errors (soup-box-scale-math.ts:316-361), and the proximity edges of a cohesive box
change (soup-box-scale-math.ts:162-215) against brute force."""
import math

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.box_scale import (CycleSchedule, computeDryBox, cyclePhaseAt, deriveCycleConfig, nextCycleTransition,
                                      proximityPairs)
from protocell_soup.rules import loadSoup


def test_schedule_reference_known_answers():
    # tests/soup-boxcycle.test.ts:162-179
    cfg = CycleSchedule(periodSteps=50000, dryFraction=0.5, cycles=5)
    assert cyclePhaseAt(0, cfg) == {"phase": "wet", "cycleIndex": 1}
    assert cyclePhaseAt(24999, cfg) == {"phase": "wet", "cycleIndex": 1}
    assert cyclePhaseAt(25000, cfg) == {"phase": "dry", "cycleIndex": 1}
    assert cyclePhaseAt(50000, cfg) == {"phase": "wet", "cycleIndex": 2}
    assert cyclePhaseAt(250000, cfg) == {"phase": "wet", "cycleIndex": 0}
    assert cyclePhaseAt(999999, cfg) == {"phase": "wet", "cycleIndex": 0}
    assert nextCycleTransition(0, cfg) == 25000
    assert nextCycleTransition(25000, cfg) == 50000
    assert nextCycleTransition(24999, cfg) == 25000
    assert nextCycleTransition(249999, cfg) == 250000
    assert nextCycleTransition(250000, cfg) == math.inf


def test_schedule_shipped_config_every_transition():
    # data/soup.json's schedule: 32,000-step periods, 65 % dry, 6 cycles -> wet 11,200 steps first
    soup = loadSoup()
    d = soup["dryWetCycle"]
    cfg = CycleSchedule(d["periodSteps"], d["dryFraction"], d["cycles"])
    ts, s = [], 0
    while True:
        s = nextCycleTransition(s, cfg)
        if s == math.inf:
            break
        ts.append(s)
    assert ts == [k * 32000 + w for k in range(6) for w in (11200, 32000)]
    for t in ts:
        assert cyclePhaseAt(t - 1, cfg)["phase"] != cyclePhaseAt(t, cfg)["phase"] or t == 6 * 32000


def test_dry_box_reference_known_answer():
    # tests/soup-boxcycle.test.ts:150-160
    box = (20.0, 20.0, 20.0)
    dry = computeDryBox(box, 3839, 0.6)
    assert 3839 / (dry[0] * dry[1] * dry[2]) == pytest.approx(0.6, abs=5e-7)
    assert dry[0] == pytest.approx(dry[1], abs=1e-10) and dry[1] == pytest.approx(dry[2], abs=1e-10)
    assert dry[0] < box[0]


def test_cycle_config_resolution_and_errors():
    soup = loadSoup()
    start = {"C": 600, "O": 100, "H": 500, "M": 20, "W": 4000}
    n = sum(start.values())
    cfg, dry = deriveCycleConfig(soup, {"dryWetCycle": True, "dryWetCycles": 2}, (20, 20, 20), n, start)
    assert cfg == CycleSchedule(32000, 0.65, 2)
    assert n / (dry[0] * dry[1] * dry[2]) == pytest.approx(soup["dryWetCycle"]["targetDryDensity"], rel=1e-12)
    assert deriveCycleConfig(soup, {}, (20, 20, 20), n, start)[0] is None
    with pytest.raises(ValueError, match="does not exceed the current broth density"):
        deriveCycleConfig(soup, {"dryWetCycle": True}, (10, 10, 10), 2000, start)
    with pytest.raises(ValueError, match="integer >= 1"):
        deriveCycleConfig(soup, {"dryWetCycle": True, "dryWetCycles": 0}, (20, 20, 20), n, start)
    # the resolved system: the list is sized for the dry density, the DH cutoff for the dry box
    opts = {"box": [20, 20, 20], "seed": 3, "kT": 1.0, "clay": False, "start": start, "dryWetCycle": True,
            "electrostatics": {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.001}}
    R = ps.resolve(opts)
    R0 = ps.resolve(dict(opts, dryWetCycle=False))
    assert R.list_cap > R0.list_cap
    assert R.cfg.es_rc <= 0.45 * min(R.dry_box) + 1e-6 < R0.cfg.es_rc
    assert (R.cycle_phase, R.cycle_index) == ("wet", 1) and (R0.cycle_phase, R0.cycle_index) == ("none", 0)
    # with solvent evaporation the dry box holds only what is left (1,223 beads): 9.7 sigma, too
    # small for the grid, which the creation-time check refuses
    with pytest.raises(ValueError, match=r"neighbour grid: box=\[9\.70"):
        ps.resolve(dict(opts, evaporateSolvent=True))


def _brute_pairs(p, box, cutoff, eligible):
    p = p.reshape(-1, 4)[:, :3].astype(np.float64)
    b = np.asarray(box, np.float64)
    out = set()
    idx = np.nonzero(eligible)[0]
    for a in idx:
        d = p[a] - p
        d = d - np.floor(d / b + 0.5) * b
        r2 = (d * d).sum(axis=1)
        for j in idx:
            if j > a and r2[j] <= cutoff * cutoff:
                out.add((int(a), int(j)))
    return out


def test_proximity_pairs_against_brute_force():
    rng = np.random.default_rng(5)
    box = (11.0, 12.5, 9.3)
    n = 400
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.random((n, 3)) * np.asarray(box)
    pos = pos.reshape(-1)
    eligible = (rng.random(n) < 0.8).astype(np.uint8)
    cut = 2.7225
    pairs = proximityPairs(pos, box, cut, eligible).reshape(-1, 2)
    got = [tuple(map(int, x)) for x in pairs]
    # at least 3 cells per axis: every pair once, i ascending, and exactly the brute-force set
    assert len(got) == len(set(got))
    assert set(got) == _brute_pairs(pos, box, cut, eligible)
    assert all(got[k][0] <= got[k + 1][0] for k in range(len(got) - 1))
    # one cell per axis: the 27 neighbour visits all land on that cell, so every pair repeats 27 times
    small = (4.0, 4.0, 4.0)
    q = np.zeros((30, 4), np.float32)
    q[:, :3] = rng.random((30, 3)) * 4.0
    q = q.reshape(-1)
    all_el = np.ones(30, np.uint8)
    p1 = [tuple(map(int, x)) for x in proximityPairs(q, small, 2.7225, all_el).reshape(-1, 2)]
    ref = _brute_pairs(q, small, 2.7225, all_el)
    assert set(p1) == ref and len(p1) == 27 * len(ref)

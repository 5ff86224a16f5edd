"""Pins the analysis oracle (CPU only) with the reference's own test cases: the recogniser cases of
tests/soup-amphiphile.test.ts:8-75, the four aggregate scenarios of tests/soup-aggregates.test.ts:
60-235 and the encapsulated-water cases of tests/water-closure.test.ts:165-273, plus the exact
closed forms of the gyration shape and the peak finder."""
import math

import numpy as np
import pytest

import analysis_scenarios as sc
from oracle import analysis as ora
from protocell_soup.analysis import loadStageThresholds

T = loadStageThresholds()
M = sc.MONOMERS


def _P(rows):
    return np.array(rows, np.float32).reshape(-1)


def test_recogniser_reference_cases():
    # tests/soup-amphiphile.test.ts:8-75 (kind 1 is the polar head O, kind 0 carbon)
    a = ora.find_amphiphiles(_P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]]), [0, 1, 1, 2, 2, 3], M)
    assert len(a) == 1 and a[0]["length"] == 3 and a[0]["chain"] == [1, 2, 3]
    assert ora.find_amphiphiles(_P([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0]]), [0, 1, 1, 2], M) == []
    assert ora.find_amphiphiles(_P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 1]]), [0, 1, 1, 2, 2, 3], M) == []
    y = ora.find_amphiphiles(_P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]]), [0, 1, 0, 2, 1, 3], M)
    assert len(y) == 1 and y[0]["headIndex"] == 0 and y[0]["length"] == 3
    assert sorted(y[0]["tailLengths"]) == [1, 2] and set(y[0]["chain"]) == {1, 2, 3}
    assert y[0]["chain"] == [1, 3, 2] and y[0]["tailLengths"] == [2, 1]  # tail 1 = first bond of the head
    t_branch = _P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0], [4, 0, 0, 0]])
    assert ora.find_amphiphiles(t_branch, [0, 1, 1, 2, 2, 3, 2, 4], M) == []
    buried = _P([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0], [4, 0, 0, 1]])
    assert ora.find_amphiphiles(buried, [0, 1, 1, 2, 2, 3, 1, 4], M) == []
    two = ora.find_amphiphiles(_P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [10, 0, 0, 1], [11, 0, 0, 0]]),
                               [0, 1, 1, 2, 3, 4], M)
    assert sorted(x["length"] for x in two) == [1, 2] and [x["headIndex"] for x in two] == [0, 3]
    # a ring (every degree 2, one head): size - 1 edges fails, rejected
    ring = _P([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0]])
    assert ora.find_amphiphiles(ring, [0, 1, 1, 2, 0, 2], M) == []


def _stage(analysis):
    ok = 1 >= T["amphiphileFraction"]
    if ok and analysis["hasVesicleAggregate"]:
        return "vesicle"
    if ok and analysis["hasLamellarAggregate"]:
        return "bilayer"
    if ok and analysis["qualifyingAggregateCount"] >= T["minMicelleAggregates"] and \
            analysis["amphiphileShareInQualifying"] >= T["minAmphiphileShareInAggregates"]:
        return "micelles"
    return "amphiphiles"


def _analyze(pos, bonds, box):
    amph = ora.find_amphiphiles(pos, bonds, M)
    mem = ora.member_indices(amph)
    return ora.analyze_aggregates(pos, box, M, amph, mem, sc.CUTOFF, T)


def test_scenario_micelles():
    a = _analyze(*sc.micelles())
    assert a["aggregateCount"] == 3 and a["sizeHistogram"] == [20, 20, 20]
    assert a["qualifyingAggregateCount"] == 3 and a["amphiphileShareInQualifying"] == 1
    assert not a["hasLamellarAggregate"] and not a["hasVesicleAggregate"] and _stage(a) == "micelles"


def test_scenario_bilayer_patch():
    a = _analyze(*sc.bilayer_patch())
    assert a["aggregateCount"] == 1
    s = a["aggregates"][0]
    assert s["flatnessRatio"] <= T["lamellarFlatnessRatio"] and s["inPlaneSymmetry"] >= T["lamellarInPlaneSymmetryMin"]
    assert s["transverseHeadShells"] == 2 and a["hasLamellarAggregate"]
    assert s["cavityVolume"] == 0 and not a["hasVesicleAggregate"] and _stage(a) == "bilayer"


def test_scenario_vesicle_shell():
    a = _analyze(*sc.vesicle_shell())
    assert a["aggregateCount"] == 1
    s = a["aggregates"][0]
    assert s["radialHeadShells"] == 2 and s["cavityVolume"] > T["enclosedVolume"]
    assert a["hasVesicleAggregate"] and not a["hasLamellarAggregate"] and _stage(a) == "vesicle"


def test_scenario_blob():
    a = _analyze(*sc.blob())
    assert a["aggregateCount"] == 1 and a["sizeHistogram"][0] > T["minAmphiphilesPerAggregate"] * 10
    assert a["qualifyingAggregateCount"] == 1 and _stage(a) not in ("micelles", "bilayer", "vesicle")


def test_shape_closed_forms():
    # a 2 x 2 x 2 cube of side 2 around (5, 6, 7): centre exact, S = I, moments 1, 1, 1, Rg sqrt(3)
    pts = [[5 + dx, 6 + dy, 7 + dz, 0] for dx in (-1, 1) for dy in (-1, 1) for dz in (-1, 1)]
    s = ora.shape_of(_P(pts))
    assert s["centre"] == [5, 6, 7] and s["principalMoments"] == [1, 1, 1] and s["radiusOfGyration"] == pytest.approx(3 ** 0.5)
    # 4 points on the (1,1,1) diagonal: S = 1.25 * ones(3, 3), eigenvalues 0, 0, 3.75
    rod = [[k, k, k, 0] for k in range(4)]
    r = ora.shape_of(_P(rod))
    assert r["principalMoments"][0] == pytest.approx(0, abs=1e-12) and r["principalMoments"][1] == pytest.approx(0, abs=1e-12)
    assert r["principalMoments"][2] == pytest.approx(3 * 1.25)
    # a flat square (z = 0): the smallest-moment axis is z
    sq = [[x, y, 0, 0] for x in (-1, 1) for y in (-1, 1)]
    q = ora.shape_of(_P(sq))
    assert q["principalMoments"][0] == 0 and [abs(v) for v in q["smallestAxis"]] == [0, 0, 1]


def test_head_peaks_reference_rules():
    bins = 10
    # two leaflets of 40 heads at z = 2.0 and 6.5 in a span of 10: separation 4.5 in [2, 10] -> 2
    rows = [[0, 0, 2.0, 1]] * 40 + [[0, 0, 6.5, 1]] * 40 + [[0, 0, 4.2, 0]] * 50
    assert ora.head_peaks(_P(rows), [1, 1, 10.0], M, bins, 2, 2, 10) == 2
    # one leaflet: no second peak beyond one sigma -> 1
    assert ora.head_peaks(_P([[0, 0, 2.0, 1]] * 80), [1, 1, 10.0], M, bins, 2, 2, 10) == 1
    # too few heads per bin -> unavailable; no heads at all -> 0
    assert ora.head_peaks(_P([[0, 0, 2.0, 1]] * 5 + [[0, 0, 7.0, 1]] * 5), [1, 1, 10.0], M, bins, 2, 2, 10) == "unavailable"
    assert ora.head_peaks(_P([[0, 0, 2.0, 0]] * 50), [1, 1, 10.0], M, bins, 2, 2, 10) == 0
    # heads in bins centred at 2.5 and 4.5: separation exactly 2.0, inside [2, 10] -> 2, outside [3, 10] -> 1
    rows = [[0, 0, 2.5, 1]] * 40 + [[0, 0, 4.0, 1]] * 40
    assert ora.head_peaks(_P(rows), [1, 1, 10.0], M, bins, 2, 2, 10) == 2
    assert ora.head_peaks(_P(rows), [1, 1, 10.0], M, bins, 2, 3, 10) == 1
    # the second bin within one sigma of the maximum is masked: one peak -> 1
    rows = [[0, 0, 2.5, 1]] * 40 + [[0, 0, 3.5, 1]] * 30
    assert ora.head_peaks(_P(rows), [1, 1, 10.0], M, bins, 2, 2, 10) == 1


def _water(pos, members, water):
    alpha = ora._json("periodic-measurement.json")["circularConcentration"]["alpha"]
    c = ora.periodic_centre(pos, members, sc.SYN_BOX, alpha)
    assert c is not None
    return ora.encapsulated_water(pos, members, water, sc.SYN_BOX, c, T["closureCell"], T["closureRadius"], T["enclosedVolume"])


def test_water_closure_equality_blob_and_hole():
    # SYN-EQUALITY: the centred and the boundary-straddling copy give the same counts
    pos, mem, wat = sc.water_shell(2001)
    res_c = _water(pos, mem, wat)
    res_s = _water(sc.shift_and_wrap(pos, [25, 25, 25]), mem, wat)
    for k in ("encapsulatedCount", "totalEmptyCells", "unreachedCells"):
        assert res_s[k] == res_c[k], k
    assert res_s["bulkWaterDensity"] == pytest.approx(res_c["bulkWaterDensity"], abs=1e-9)
    assert 900 * 0.9 < res_c["encapsulatedCount"] < 900 * 1.1 and res_c["closed"]
    # SYN-BLOB: a solid ball traps nothing
    pos, mem, wat = sc.water_blob(2002)
    rb = _water(pos, mem, wat)
    assert rb["encapsulatedCount"] == 0 and not rb["closed"]
    # SYN-HOLE: a 30-degree full-thickness hole lets the water out
    pos, mem, wat = sc.water_shell(2003)
    closed = _water(pos, mem, wat)
    hole_dir = [1 / math.sqrt(3)] * 3
    pos, mem, wat = sc.water_shell(2003, hole=(hole_dir, 30.0))
    holed = _water(pos, mem, wat)
    assert closed["encapsulatedCount"] > 900 * 0.9
    assert holed["encapsulatedCount"] < closed["encapsulatedCount"] * 0.05 and not holed["closed"]

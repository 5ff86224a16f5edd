"""GPU parity of the snapshot analysis (soup_analysis_* on the device plus the host geometry of
protocell_soup.analysis) against the oracle (oracle/analysis.py over analysis_oracle.c): amphiphile
lists, member order, aggregate labels and sizes and encapsulated-water counts bit-exact; the FP64
per-aggregate geometry to 1e-9 relative."""
import math

import numpy as np
import pytest

import analysis_scenarios as sc
import protocell_soup as ps
from oracle import analysis as ora
from protocell_soup import analysis as an
from soup_fixtures import amphiphile_opts

pytestmark = pytest.mark.gpu

T = an.loadStageThresholds()
M = sc.MONOMERS


@pytest.fixture(scope="module")
def ctx():
    c = an.AnalysisContext()
    yield c
    c.close()


def _close(a, b, key):
    if isinstance(a, (list, tuple)):
        assert len(a) == len(b), key
        for x, y in zip(a, b):
            _close(x, y, key)
    elif isinstance(a, float) or isinstance(b, float):
        if math.isnan(a) and math.isnan(b):
            return
        assert a == pytest.approx(b, rel=1e-9, abs=1e-9), key
    else:
        assert a == b, key


def _same_analysis(g, o):
    for k in ("aggregateCount", "sizeHistogram", "qualifyingAggregateCount", "amphiphilesInQualifying",
              "hasLamellarAggregate", "hasVesicleAggregate"):
        assert g[k] == o[k], k
    _close(g["amphiphileShareInQualifying"], o["amphiphileShareInQualifying"], "share")
    assert len(g["aggregates"]) == len(o["aggregates"])
    for sg, so in zip(g["aggregates"], o["aggregates"]):
        assert list(sg) == list(so)
        for k in sg:
            if k == "encapsulatedWater":
                if so[k] is None:
                    assert sg[k] is None
                    continue
                for kk in so[k]:
                    _close(sg[k][kk], so[k][kk], kk)
            else:
                _close(sg[k], so[k], k)


def test_recogniser_matches_oracle_on_reference_cases(ctx):
    cases = [
        ([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]], [0, 1, 1, 2, 2, 3]),
        ([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]], [0, 1, 0, 2, 1, 3]),
        ([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0], [4, 0, 0, 0]], [0, 1, 1, 2, 2, 3, 2, 4]),
        ([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0], [4, 0, 0, 1]], [0, 1, 1, 2, 2, 3, 1, 4]),
        ([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0], [10, 0, 0, 1], [11, 0, 0, 0]], [0, 1, 1, 2, 3, 4]),
        ([[0, 0, 0, 1], [1, 0, 0, 0], [2, 0, 0, 0]], [0, 1, 1, 2, 0, 2]),
    ]
    for rows, bonds in cases:
        pos = np.array(rows, np.float32).reshape(-1)
        assert an.findAmphiphiles(pos, np.array(bonds), M, ctx) == ora.find_amphiphiles(pos, bonds, M)


@pytest.mark.parametrize("scenario", ["micelles", "bilayer_patch", "vesicle_shell", "blob"])
def test_aggregate_scenarios_match_oracle(ctx, scenario):
    pos, bonds, box = getattr(sc, scenario)()
    amph = an.findAmphiphiles(pos, bonds, M, ctx)
    o_amph = ora.find_amphiphiles(pos, bonds, M)
    assert amph == o_amph
    mem = an.memberIndicesOf(amph)
    assert mem == ora.member_indices(o_amph)
    g = an.analyzeAggregates(ctx, pos, box, M, amph, mem, sc.CUTOFF, T)
    o = ora.analyze_aggregates(pos, box, M, o_amph, mem, sc.CUTOFF, T)
    _same_analysis(g, o)


def test_water_closure_matches_oracle(ctx):
    alpha = an.circularAlpha()
    for pos, mem, wat in (sc.water_shell(2001), sc.water_blob(2002), sc.water_shell(2003, hole=([3 ** -0.5] * 3, 30.0))):
        for p in (pos, sc.shift_and_wrap(pos, [25, 25, 25])):
            ctx.load(p, np.full(3 * (p.size // 4), 0xFFFFFFFF, np.uint32))
            centre = an.periodicCentreOf(p, mem, sc.SYN_BOX, alpha)
            o_centre = ora.periodic_centre(p, mem, sc.SYN_BOX, alpha)
            _close(centre, o_centre, "centre")
            g = an.encapsulatedWaterVolume(ctx, mem, 1, sc.SYN_BOX, o_centre, T["closureCell"], T["closureRadius"],
                                           T["enclosedVolume"])
            o = ora.encapsulated_water(p, mem, wat, sc.SYN_BOX, o_centre, T["closureCell"], T["closureRadius"],
                                       T["enclosedVolume"])
            for k in o:
                _close(g[k], o[k], k)


def test_detect_stage_on_a_live_soup_matches_oracle(ctx):
    opts = amphiphile_opts(18.0, 220, 3, {"W": 2200, "C": 150, "M": 10}, seed=4,
                           electrostatics={"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01})
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(2000)
        stage, ev = an.detectStage(s, ctx)
        pos = s.particles()
        bonds = s.bonds()
        o_amph = ora.find_amphiphiles(pos, bonds, M)
        assert [a["headIndex"] for a in o_amph] and len(o_amph) >= 200
        mem = ora.member_indices(o_amph)
        water_kind = next(k for k, m in enumerate(M) if m["id"] == "W")
        water = list(np.nonzero(np.rint(pos.reshape(-1, 4)[:, 3]) == water_kind)[0])
        o = ora.analyze_aggregates(pos, list(s.box), M, o_amph, mem, an.memberCutoff(), T, water)
        _same_analysis(ev["aggregateAnalysis"], o)
        assert ev["enclosedVolume"] == pytest.approx(
            ora.enclosed_volume_box(pos, mem, list(s.box), T["closureCell"], T["closureRadius"]), abs=0)
        carbon = [k for k, m in enumerate(M) if m["kind"] == "carbon"]
        n_c = int(np.isin(np.rint(pos.reshape(-1, 4)[:, 3]), carbon).sum())
        assert ev["amphiphileFraction"] == sum(a["length"] for a in o_amph) / n_c
        comps = ora.cluster_components(pos.reshape(-1, 4)[mem], list(s.box), an.memberCutoff(), True)
        assert ev["largestAggregateFraction"] == len(comps[0]) / len(mem)
        assert ev["headPeaks"] == ora.head_peaks(pos, list(s.box), M, T["headDensityBins"], T["minHeadsPerBin"],
                                                 T["headPeakSeparationMin"], T["headPeakSeparationMax"])
        assert stage in ("amphiphiles", "micelles", "bilayer", "vesicle")
    finally:
        s.dispose()

"""The campaign's chemistry at full size against the reference's own box-76 campaign trace.

tests/golden/bbB76_wet_trace.json holds the wet rows (steps 3,000, 8,000 and 11,000, before the first
drying ramp) of verify/out/big-box-campaign-B76-trace.json: the reference's campaign with the
configuration this build benchmarks (483,268 beads in a 76 sigma box, seed 19, kT 1.1, pH 7.0,
0.01 M, --relax --cycle --evaporate --cycles 1). The same campaign is run here on the GPU through the
campaign CLI's own configuration path, with one difference the reference's history requires: that
campaign ran before the acid-soap head pair existed (its later report reads "the bbB76 checkpoints
already on disk ... this task did not run the campaign", acid-soap-pairing-report.md:388-391), so the
pair is off here (acidSoapScaleOverride = 0), in the dynamics and in the constant-pH work alike. The
long-range DH list and the catalyst-turnover fix both predate the campaign and stay on.

Trajectories of the two implementations part at the f32 level within a few hundred steps, so the
comparison is statistical: each cumulative count must lie within
max(10 %, 3 sqrt(count)) of the reference's, and the valence invariants must read zero as they do
there. A chemistry restated wrongly in both the kernels and the oracle (a nucleation tip on the wrong
carbon, a missing interval factor in the attempt probability, a head allowed mid-chain) moves these
counts far outside that band.
"""
import json
import math
import os

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.analysis import findAmphiphiles
from protocell_soup.campaign import parseCliArgs, runConfig
from protocell_soup.rules import loadSoup

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TRACE = json.load(open(os.path.join(HERE, "golden", "bbB76_wet_trace.json")))
FLAGS = ["--label", "bbB76", "--box", "76", "--start", '{"C":55988,"O":18659,"H":55988,"M":1452,"W":351181}', "--seed", "19",
         "--kT", "1.1", "--relax", "--cycle", "--evaporate", "--cycles", "1", "--charge", "--pH", "7.0", "--ionicStrength", "0.01",
         "--steps", "11000"]


def _close(ours, ref):
    return abs(ours - ref) <= max(0.10 * ref, 3.0 * math.sqrt(max(ref, 1)))


def _valence(pos, edges, soup):
    """The trace's invariants.valence counters (soup/cli/campaign.ts's invariant report)."""
    kinds = np.rint(pos.reshape(-1, 4)[:, 3]).astype(np.int64)
    kind_of = [m["kind"] for m in soup["monomers"]]
    n = kinds.size
    cc, co, hd, deg = (np.zeros(n, int) for _ in range(4))
    bondable = {"carbon", "head"}
    nonbondable = 0
    for i, j in edges:
        ki, kj = kind_of[kinds[i]], kind_of[kinds[j]]
        deg[i] += 1
        deg[j] += 1
        if ki not in bondable or kj not in bondable:
            nonbondable += 1
        if ki == kj == "carbon":
            cc[i] += 1
            cc[j] += 1
        elif ki == "carbon" and kj == "head":
            co[i] += 1
            hd[j] += 1
        elif ki == "head" and kj == "carbon":
            co[j] += 1
            hd[i] += 1
    carbon = np.array([kind_of[k] == "carbon" for k in kinds])
    head = np.array([kind_of[k] == "head" for k in kinds])
    cap = soup["headPlacement"]["chainCapacity"]
    # a head's carbon must be a chain end: no other C-C neighbour beyond one
    return {"carbonCCover2": int((carbon & (cc > 2)).sum()), "carbonCOover1": int((carbon & (co > 1)).sum()),
            "headOverChainCapacity": int((head & (hd > cap)).sum()), "nonBondableBonded": nonbondable,
            "degreeOver3": int((deg > 3).sum()), "headsWithTwoTails": int((head & (hd == 2)).sum())}


def test_wet_campaign_chemistry_matches_the_reference_trace():
    soup = loadSoup()
    cfg = runConfig(parseCliArgs(FLAGS))
    opts = {k: cfg[k] for k in ("box", "seed", "kT", "start", "dryWetCycle", "dryWetCycles", "evaporateSolvent", "electrostatics")
            if k in cfg}
    opts["clay"] = False
    opts["acidSoapScaleOverride"] = 0
    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        for row in TRACE["rows"]:
            s.stepCycled(row["step"] - s.steps)
            assert s.steps == row["step"] and s.N == row["N"] and s.box == [76.0, 76.0, 76.0]
            ev = s.events()
            pos = s.particles()
            edges = s.bonds().reshape(-1, 2).astype(np.int64)
            amph = len(findAmphiphiles(pos, s.bonds(), soup["monomers"]))
            val = _valence(pos, edges, soup)
            print(f"CAMPAIGN step={row['step']} cc_bond={ev['cc_bond']} ({row['events']['cc_bond']}) "
                  f"co_bond={ev['co_bond']} ({row['events']['co_bond']}) co_break={ev['co_break']} ({row['events']['co_break']}) "
                  f"bonds={len(edges)} ({row['bonds']}) amphiphiles={amph} ({row['amphiphileCount']}) "
                  f"twoTailed={val['headsWithTwoTails']} ({row['valence']['headsWithTwoTails']})", flush=True)
            assert _close(ev["cc_bond"], row["events"]["cc_bond"])
            assert _close(ev["co_bond"], row["events"]["co_bond"])
            assert ev["cc_break"] == row["events"]["cc_break"] == 0
            assert _close(ev["co_break"], row["events"]["co_break"])
            assert _close(len(edges), row["bonds"])
            assert _close(amph, row["amphiphileCount"])
            assert _close(val["headsWithTwoTails"], row["valence"]["headsWithTwoTails"])
            for k in ("carbonCCover2", "carbonCOover1", "headOverChainCapacity", "nonBondableBonded", "degreeOver3"):
                assert val[k] == row["valence"][k] == 0, k
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
    finally:
        s.dispose()

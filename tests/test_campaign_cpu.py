"""The campaign CLI's host side (no GPU): flags and defaults (soup/cli/campaign-config.ts:187-246),
the resume signature's exact text (:252-283), the newest-matching-checkpoint scan (:285-309), the
refusal of parcel confinement, and the acid-soap pairing statistics (soup/src/electrostatics.ts
:520-565) against a brute-force loop over every head pair."""
import json
import math
import os

import numpy as np
import pytest

from protocell_soup.campaign import configSignature, findNewestMatchingCheckpoint, main, parseCliArgs, run, runConfig
from protocell_soup.checkpoint import jsonStringify, writeCheckpointFile
from protocell_soup.electrostatics import makeEsBasis, pairingStats
from protocell_soup.rules import loadParams, loadSoup


def test_flags_defaults_and_required():
    soup, p = loadSoup(), loadParams()
    a = parseCliArgs(["--label", "x", "--steps", "500", "--box", "30", "--start", '{"W":100,"C":300}'])
    assert a["every"] == soup["checkpoint"]["everySteps"] and a["dir"] == soup["checkpoint"]["dir"]
    assert a["seed"] == 1 and a["kT"] == p["thermostat"]["kT"] and a["start"] == {"W": 100, "C": 300}
    assert a["expandRampSteps"] == 15 and a["expandRampRelaxSteps"] == 60 and a["confineStiffness"] == 100
    assert a["minimiseIterations"] == soup["coldStartRelax"]["iterations"] and a["minimiseAt"] == []
    assert not (a["cycle"] or a["evaporate"] or a["relax"] or a["charge"])
    b = parseCliArgs(["--label", "x", "--steps", "5", "--box", "20", "--start", "{}", "--minimiseAt", "900, 100,500", "--cycle"])
    assert b["minimiseAt"] == [100, 500, 900] and b["cycle"]
    with pytest.raises(ValueError, match="required flags missing: --steps, --box"):
        parseCliArgs(["--label", "x", "--start", "{}"])
    assert parseCliArgs(["--help"]) == {"help": True}
    assert main(["--label", "x"]) == 1


def test_signature_text_is_the_reference_one():
    a = parseCliArgs(["--label", "x", "--steps", "500", "--box", "30", "--start", '{"W":100,"C":300}'])
    cfg = runConfig(a)
    # kT defaults to data/params.json's thermostat.kT (1.1)
    assert jsonStringify(cfg) == '{"box":[30,30,30],"seed":1,"kT":1.1,"start":{"W":100,"C":300}}'
    # the start keys sorted, the three plain fields as null, the conditional ones absent
    assert configSignature(cfg) == ('{"box":[30,30,30],"seed":1,"kT":1.1,"start":{"C":300,"W":100},"catalystCount":null,'
                                    '"dryWetCycle":null,"evaporateSolvent":null}')
    b = parseCliArgs(["--label", "x", "--steps", "1", "--box", "24.5", "--start", '{"O":3,"C":2}', "--seed", "7", "--kT", "1.1",
                      "--catalyst", "12", "--cycle", "--evaporate", "--cycles", "1", "--minimiseAt", "500,100", "--charge",
                      "--pH", "6"])
    assert configSignature(runConfig(b)) == (
        '{"box":[24.5,24.5,24.5],"seed":7,"kT":1.1,"start":{"C":2,"O":3},"catalystCount":12,"dryWetCycle":true,'
        '"evaporateSolvent":true,"dryWetCycles":1,"minimiseAt":[100,500],'
        '"electrostatics":{"enabled":true,"pH":6,"ionicStrengthMolar":0.1}}')


def test_newest_matching_checkpoint(tmp_path):
    d = str(tmp_path)
    cfg_a = {"box": [30, 30, 30], "seed": 1, "kT": 1, "start": {"C": 300}}
    cfg_b = dict(cfg_a, seed=2)
    for step, cfg, label in ((100, cfg_a, "run"), (300, cfg_a, "run"), (500, cfg_b, "run"), (999, cfg_a, "other")):
        writeCheckpointFile(d, label, {"version": 1, "config": cfg, "globalStep": step})
    with open(os.path.join(d, "run-step700.json"), "w") as fh:
        fh.write('{"version": 1, "config": ')  # killed mid-write
    with open(os.path.join(d, "run-step900.json.tmp"), "w") as fh:
        fh.write(json.dumps({"version": 1, "config": cfg_a, "globalStep": 900}))
    found = findNewestMatchingCheckpoint(d, "run", configSignature(cfg_a))
    assert found["file"]["globalStep"] == 300 and found["path"].endswith("run-step300.json")
    assert findNewestMatchingCheckpoint(d, "run", configSignature(cfg_b))["file"]["globalStep"] == 500
    assert findNewestMatchingCheckpoint(d, "none", configSignature(cfg_a)) is None
    assert findNewestMatchingCheckpoint(os.path.join(d, "missing"), "run", configSignature(cfg_a)) is None


def test_confinement_is_refused(tmp_path):
    a = parseCliArgs(["--label", "x", "--steps", "5", "--box", "20", "--start", "{}", "--dir", str(tmp_path),
                      "--confineRadius", "8"])
    with pytest.raises(ValueError, match="confinement"):
        run(a, out=lambda s: None)


def _brute_pairing(pos, q, box, kind, contact):
    p = pos.reshape(-1, 4)
    heads = [i for i in range(p.shape[0]) if math.trunc(float(p[i, 3])) == kind]
    unlike = like = withn = paired = charged = 0
    for i in heads:
        charged += q[i] != 0
        best, bj, anyn = math.inf, -1, False
        for j in heads:
            if j == i:
                continue
            d2 = 0.0
            for a in range(3):
                d = float(p[i, a]) - float(p[j, a])
                x = d / box[a]
                r = math.floor(x)
                m = d - ((r + 1) if x - r >= 0.5 else r) * box[a]
                d2 += m * m
            r = math.sqrt(d2)
            if r > contact:
                continue
            anyn = True
            if (q[i] == 0) != (q[j] == 0):
                unlike += 1
            else:
                like += 1
            if r < best:
                best, bj = r, j
        withn += anyn
        if bj >= 0 and (q[i] == 0) != (q[bj] == 0):
            paired += 1
    return {"heads": len(heads), "withNeighbour": withn, "unlikeContacts": unlike, "likeContacts": like, "paired": paired,
            "alpha": charged / len(heads)}


def test_pairing_stats_against_brute_force():
    soup, p = loadSoup(), loadParams()
    es = makeEsBasis(soup, p, {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1})
    rng = np.random.default_rng(4)
    box = [9.0, 11.0, 10.0]
    n = 700
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (rng.random((n, 3)) * np.asarray(box)).astype(np.float32)
    pos[:, 3] = rng.choice([0, es.chargedKind, 4], size=n, p=[0.3, 0.5, 0.2])
    q = np.where((pos[:, 3] == es.chargedKind) & (rng.random(n) < 0.4), np.float32(-1), np.float32(0)).astype(np.float32)
    for contact in (1.0, 1.6):
        got = pairingStats(pos.reshape(-1), q, box, es, contact)
        exp = _brute_pairing(pos, q, box, es.chargedKind, contact)
        for k, v in exp.items():
            assert got[k] == v, (contact, k)
        assert got["unlikeContacts"] > 0 and got["likeContacts"] > 0 and got["paired"] > 0
        assert got["unlikeFractionRandom"] == 2 * got["alpha"] * (1 - got["alpha"])

"""The campaign CLI on the GPU (soup/cli/campaign.ts): a fresh run writes reference-layout
checkpoints every block and a progress line per block; the next call with the same flags resumes
from the newest matching checkpoint and ends in exactly the state of one uninterrupted run of the
same blocks; --minimiseAt cuts a block and minimises at that step; --expandTo grows the box once
(no particle moved) and checkpoints it, and a resume does not expand again."""
import json
import os

import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.campaign import parseCliArgs, run, runConfig
from protocell_soup.checkpoint import decodeFloat32, decodeUint32, readCheckpointFile

pytestmark = pytest.mark.gpu

START = {"C": 600, "O": 100, "H": 500, "M": 20, "W": 2000}


def _argv(d, label, steps, *extra, start=START):
    return ["--label", label, "--steps", str(steps), "--every", "100", "--dir", str(d), "--box", "16", "--start",
            json.dumps(start), "--seed", "5", "--kT", "1", *extra]


def test_fresh_run_then_resume_equals_one_uninterrupted_run(tmp_path):
    lines = []
    args = parseCliArgs(_argv(tmp_path, "t", 300, "--relax"))
    r = run(args, out=lines.append)
    print("\n".join(lines), flush=True)
    assert r == {"step": 300, "stopped": False, "saved": os.path.join(str(tmp_path), "t-step300.json")}
    assert sorted(os.listdir(tmp_path)) == ["t-step100.json", "t-step200.json", "t-step300.json"]
    assert lines[0].startswith("[campaign] new run label=t")
    assert any(s.startswith("[campaign] cold-start minimisation: iterations=200") for s in lines)
    prog = [s for s in lines if s.startswith("[campaign] step=")]
    assert [s.split()[1] for s in prog] == ["step=100/300", "step=200/300", "step=300/300"]
    assert all(" stage=" in s and " census=" in s and " phase=none/0 " in s for s in prog)
    f = readCheckpointFile(os.path.join(str(tmp_path), "t-step300.json"))
    assert f["version"] == 1 and f["globalStep"] == 300 and f["config"] == runConfig(args) and f["N"] == 3220

    lines2 = []
    r2 = run(parseCliArgs(_argv(tmp_path, "t", 200, "--relax")), out=lines2.append)
    assert lines2[0] == f"[campaign] resume label=t from {os.path.join(str(tmp_path), 't-step300.json')}, step=300"
    assert any("--relax skipped: this is a resume from step=300" in s for s in lines2)
    assert r2["step"] == 500 and lines2[-1] == "[campaign] step budget fully completed: step=500"

    s = ps.createSoup({"box": [16, 16, 16], "seed": 5, "kT": 1, "start": START, "clay": False})
    try:
        s.relaxColdStart()
        for _ in range(5):
            s.stepCycled(100)
        g = readCheckpointFile(os.path.join(str(tmp_path), "t-step500.json"))
        assert np.array_equal(decodeFloat32(g["positionsB64"]), s.particles())
        assert np.array_equal(decodeFloat32(g["velocitiesB64"]), s.velocities())
        assert np.array_equal(decodeUint32(g["bondSlotsB64"]), s.bondSlots())
        assert np.array_equal(decodeUint32(g["thermoRngB64"]), s.rngState()["thermo"])
        assert g["events"] == s.events()
    finally:
        s.dispose()


def test_minimise_at_cuts_a_block(tmp_path):
    lines = []
    r = run(parseCliArgs(_argv(tmp_path, "m", 200, "--relax", "--minimiseAt", "150", "--minimiseIterations", "10")),
            out=lines.append)
    assert r["step"] == 200
    assert sorted(os.listdir(tmp_path)) == ["m-step100.json", "m-step150.json", "m-step200.json"]
    assert any(s.startswith("[campaign] mid-run minimisation at step=150 iterations=10 ") for s in lines)
    f = readCheckpointFile(os.path.join(str(tmp_path), "m-step200.json"))
    assert f["config"]["minimiseAt"] == [150]


def test_expand_to_grows_once_and_checkpoints(tmp_path):
    water = {"C": 0, "O": 0, "H": 0, "M": 0, "W": 2000}  # nothing bonded across the boundary the growth opens
    argv = _argv(tmp_path, "e", 100, "--relax", "--expandTo", "17", "--expandRampSteps", "3", "--expandRampRelaxSteps", "0",
                 start=water)
    lines = []
    r = run(parseCliArgs(argv), out=lines.append)
    print("\n".join(lines), flush=True)
    assert r["step"] == 100
    assert any(s.startswith("[campaign] invariant confirmed") for s in lines)
    assert any(s.startswith("[campaign] before expansion (probe): box=[16,16,16] steps=0 n=100") for s in lines)
    assert any(s.startswith("[campaign] after expansion (") and "box=[17,17,17] steps=0" in s for s in lines)
    assert sorted(os.listdir(tmp_path)) == ["e-step0.json", "e-step100.json"]
    f = readCheckpointFile(os.path.join(str(tmp_path), "e-step100.json"))
    assert f["liveBox"] == [17, 17, 17]
    lines2 = []
    run(parseCliArgs(argv), out=lines2.append)
    assert any(s == "[campaign] box already expanded to [17,17,17] -- skipping (idempotent resume)" for s in lines2)


def test_expand_to_with_ramp_dynamics_fails_the_fingerprint_check(tmp_path):
    # the reference's check (campaign.ts:309-321) asserts that no particle moved across growBoxTo,
    # whose ramp runs rampRelaxSteps of ordinary dynamics between increments: with any such steps
    # the fingerprint changes and the run stops there, in the reference and here alike
    water = {"C": 0, "O": 0, "H": 0, "M": 0, "W": 2000}
    argv = _argv(tmp_path, "f", 100, "--expandTo", "17", "--expandRampSteps", "3", "--expandRampRelaxSteps", "10", start=water)
    with pytest.raises(RuntimeError, match="moved at least one particle"):
        run(parseCliArgs(argv), out=lambda s: None)
    assert os.listdir(tmp_path) == []

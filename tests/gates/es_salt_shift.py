"""Long-run gate: the salt shift of the apparent pKa on the reference's charged campaign.

The reference calibrated its long-range electrostatics (tests/es-range-calibration.test.ts) on two
checkpoints of its charged box-54 campaign, run as

    campaign.ts --label esB54pH50 --box 54 --start '{"C":27894,"O":9296,"H":27894,"M":723,"W":125971}'
      --seed 19 --kT 1.1 --relax --cycle --evaporate --cycles 1 --charge --pH 5.0

one in the dry phase (step 31,000, box 36.637, the solvent evaporated) and one wet (step 174,400,
after rehydration). On each frozen configuration it re-equilibrated the protonation Monte Carlo at
pH 6.0 for 16 sweeps from PCG state 777 at ionic strengths 0.01 and 0.1 M, with the shipped
cutoff (4 Debye lengths, capped at 0.45 of the box), averaged alpha over the last 8 sweeps and read
the apparent pKa by inverse Henderson-Hasselbalch (verify/out/es-range-calibration-{dry,wet}-pH6.json).

This script runs the same campaign on this library to the same two global steps and writes the two
checkpoints, then the same calibration on them. Three settings differ from the flags above, all
because the reference's checkpoints and calibration predate two later features of its own model
(the calibration's own header: "the configuration was generated at the OLD range"):
- the acid-soap head pair (data/soup.json's acidSoapPair, added after the calibration) is off in the
  dynamics and in the calibration (acidSoapScaleOverride = 0). With it on, the same calibration
  reads alpha 0.36-0.56 (pKa 5.90-6.25, salt shift 0.25-0.26): the pair term pulls alpha toward
  one half, as its own basis says;
- the dynamics use the predecessor's screened-Coulomb cutoff, 2.7224620 sigma (cutoffSigma);
- the head list is sized for heads gathered into an eighth of the dry box (densityVolumeSigma3):
  with the shipped long-range rule the first run of this campaign overflowed the uniform-density
  capacity (318) in the evaporating ramp, a loud stop, as the reference's own capacity rule
  predicts once the heads aggregate.
The trajectories are independent realisations (f32 dynamics diverge), so the comparison is
statistical: the apparent pKa of each arm and the salt shift between 0.01 and 0.1 M, against the
reference's numbers.

Run on a GPU: python tests/gates/es_salt_shift.py out.json [checkpoint-dir]
"""
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "protocell-genesis_amd"))

import numpy as np  # noqa: E402

import protocell_soup as ps  # noqa: E402
from protocell_soup.box_scale import nextCycleTransition  # noqa: E402
from protocell_soup.checkpoint import decodeCheckpointResume, encodeCheckpoint, readCheckpointFile, writeCheckpointFile  # noqa: E402
from protocell_soup.electrostatics import apparentPKa  # noqa: E402

START = {"C": 27894, "O": 9296, "H": 27894, "M": 723, "W": 125971}
CONFIG = {"box": [54, 54, 54], "seed": 19, "kT": 1.1, "start": START, "dryWetCycle": True, "evaporateSolvent": True,
          "dryWetCycles": 1, "electrostatics": {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1},
          "acidSoapScaleOverride": 0}
DYNAMICS_ES = {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1, "cutoffSigma": 2.722462,
               "densityVolumeSigma3": 36.637031822882015 ** 3 / 8}
# the shipped-cutoff rows of verify/out/es-range-calibration-{dry,wet}-pH6.json
REFERENCE = {
    "dry": {"step": 31000, "box": 36.637031822882015, "pKa": {0.01: 6.748195, 0.1: 5.941503}},
    "wet": {"step": 174400, "box": 54.0, "pKa": {0.01: 6.120088, 0.1: 5.115746}},
}
PH, SWEEPS, RNG0 = 6.0, 16, 777


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def calibrate(file, ionic):
    """16 protonation sweeps on the frozen configuration at pH 6.0 and `ionic` M, from PCG state 777
    and the checkpoint's own charges; alpha over the last 8."""
    cfg = file["config"]
    resume = decodeCheckpointResume(file)
    resume["protonationRng"] = RNG0
    opts = {"box": cfg["box"], "seed": cfg["seed"], "kT": cfg["kT"], "start": cfg["start"], "clay": False, "resume": resume,
            "acidSoapScaleOverride": 0,
            # the head list sized for aggregated heads (the reference's CPU sweep has no list to overflow)
            "electrostatics": {"enabled": True, "pH": PH, "ionicStrengthMolar": ionic,
                               "densityVolumeSigma3": float(np.prod(file["liveBox"])) / 8}}
    s = ps.createSoup(opts)
    try:
        traj = [s.protonationSweepDEBUG()["alpha"] for _ in range(SWEEPS)]
        es = s.electrostatics()
    finally:
        s.dispose()
    half = traj[SWEEPS // 2:]
    alpha = float(np.mean(half))
    return {"I": ionic, "alpha": alpha, "alphaSd": float(np.sqrt(np.mean((np.asarray(half) - alpha) ** 2))),
            "pKaApp": apparentPKa(alpha, PH), "cutoff": es["cutoff"], "debyeLengths": es["debyeLengthsSpanned"],
            "alphaTrajectory": traj}


def advance_to(s, target):
    """stepCycled to global step `target`, stopping at each schedule transition first (a transition
    adds its ramp's own steps, which a single call would carry past the target)."""
    while s.steps < target:
        nxt = nextCycleTransition(s.steps, s._R.cycle_cfg)
        s.stepCycled(int(min(target, nxt, s.steps + 20000) - s.steps))
        log(f"[es-salt] step={s.steps} phase={s.cyclePhase} box={s.box[0]:.4f} N={s.N} nonFinite={s.nonFiniteCount()}")


def main():
    out_path = sys.argv[1] if len(sys.argv) > 1 else None
    ck_dir = sys.argv[2] if len(sys.argv) > 2 else "/tmp/es_salt_shift_checkpoints"
    os.makedirs(ck_dir, exist_ok=True)
    t0 = time.time()
    s = ps.createSoup(dict(CONFIG, electrostatics=DYNAMICS_ES, clay=False))
    try:
        r = s.relaxColdStart()
        log(f"[es-salt] N={s.N} relaxed max|F| {r['maxForceBefore']:.4g} -> {r['maxForceAfter']:.4g}")
        for ref in (REFERENCE["dry"], REFERENCE["wet"]):
            advance_to(s, ref["step"])
            p = writeCheckpointFile(ck_dir, "esB54pH50", encodeCheckpoint(s, CONFIG))
            log(f"[es-salt] checkpoint {p} box={s.box} N={s.N}")
    finally:
        s.dispose()
    campaign_s = time.time() - t0
    result = {"campaign": CONFIG, "dynamicsElectrostatics": DYNAMICS_ES, "campaignSeconds": campaign_s, "arms": {}}
    for arm, ref in REFERENCE.items():
        f = readCheckpointFile(os.path.join(ck_dir, f"esB54pH50-step{ref['step']}.json"))
        rows = [calibrate(f, I) for I in (0.01, 0.1)]
        for r in rows:
            r["referencePKaApp"] = ref["pKa"][r["I"]]
            log(f"[es-salt] {arm} step={f['globalStep']} box={f['liveBox'][0]:.4f} N={f['N']} I={r['I']} "
                f"rc={r['cutoff']:.3f} ({r['debyeLengths']:.2f} lD) alpha={r['alpha']:.4f}+-{r['alphaSd']:.4f} "
                f"pKaApp={r['pKaApp']:.4f} (reference {r['referencePKaApp']:.4f})")
        shift = rows[0]["pKaApp"] - rows[1]["pKaApp"]
        ref_shift = ref["pKa"][0.01] - ref["pKa"][0.1]
        result["arms"][arm] = {"step": f["globalStep"], "box": f["liveBox"], "N": f["N"], "activeCounts": f["activeCounts"],
                               "rows": rows, "saltShift": shift, "referenceSaltShift": ref_shift}
        log(f"[es-salt] {arm}: salt shift pKa(0.01) - pKa(0.1) = {shift:.4f} (reference {ref_shift:.4f})")
    text = json.dumps(result, indent=1)
    print(text)
    if out_path:
        with open(out_path, "w") as fh:
            fh.write(text)


if __name__ == "__main__":
    main()

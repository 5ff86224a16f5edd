"""Long-run gate: the live salt shift of the apparent pKa at box 30 (the reference's +0.709).

The reference measured its one independent validation of the long-range electrostatics on two live
runs of its box-30 broth, identical but for the ionic strength
(.superpowers/sdd/2026-08-16-soup-to-vesicle/long-range-electrostatics-report.md:398-414, 495;
docs/verdict.md:123-126):

    campaign.ts --box 30 --start '{"C":4783,"O":1594,"H":4783,"M":124,"W":21600}' --seed 19 --kT 1.1
      --relax --cycle --evaporate --charge --pH 6.0 --ionicStrength 0.01   (and 0.1)

read at step 40,000 (the second wet segment): alpha 0.6192, pKa_app 5.789 at 0.01 M (cutoff 9.1592
sigma, 2.410 Debye lengths, capped by the dry box); alpha 0.8927, pKa_app 5.080 at 0.1 M (4.8067
sigma, 4 Debye lengths): salt shift +0.709 against the literature's ~0.7.

This script runs both arms on this library for seed 19 and further seeds. Two settings follow from
the date of that measurement:
- the acid-soap head pair (data/soup.json's acidSoapPair) came later, so it is off
  (acidSoapScaleOverride = 0); a second pass with the shipped pair on is reported beside it;
- the head list is sized for heads gathered into an eighth of the dry box (densityVolumeSigma3): the
  reference ran these arms with a fixed capacity of 2,500 (report:384-386).
The trajectories are independent realisations (the f32 dynamics diverge), so the comparison is
statistical. Each arm records alpha after every constant-pH sweep (every 1,000 steps); the value at
step 40,000 is the reference's quantity, and the mean over the sweeps at 36,000-40,000 is reported
beside it.

Run on a GPU: python tests/gates/es_salt_shift_live.py out.json [seed ...]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "protocell-genesis_amd"))

import numpy as np  # noqa: E402

import protocell_soup as ps  # noqa: E402
from protocell_soup.box_scale import nextCycleTransition  # noqa: E402
from protocell_soup.electrostatics import apparentPKa  # noqa: E402

START = {"C": 4783, "O": 1594, "H": 4783, "M": 124, "W": 21600}
BOX, PH, TARGET, READ_FROM = 30.0, 6.0, 40000, 36000
REFERENCE = {"seed": 19, "step": TARGET, "alpha": {0.01: 0.6192, 0.1: 0.8927}, "pKaApp": {0.01: 5.789, 0.1: 5.080},
             "cutoff": {0.01: 9.1592, 0.1: 4.8067}, "saltShift": 0.709}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def dry_box_volume():
    """The dry box of the evaporating cycle at creation (the head list is sized for an eighth of it)."""
    return float(np.prod(ps.resolve(dict(config(19, 0.01, 0.0, None), clay=False)).dry_box))


def config(seed, ionic, acid_soap, vol):
    es = {"enabled": True, "pH": PH, "ionicStrengthMolar": ionic}
    if vol is not None:
        es["densityVolumeSigma3"] = vol / 8
    opts = {"box": [BOX, BOX, BOX], "seed": seed, "kT": 1.1, "start": START, "dryWetCycle": True, "evaporateSolvent": True,
            "electrostatics": es}
    if acid_soap is not None:
        opts["acidSoapScaleOverride"] = acid_soap
    return opts


def run_arm(seed, ionic, acid_soap, vol):
    t0 = time.time()
    s = ps.createSoup(dict(config(seed, ionic, acid_soap, vol), clay=False))
    trace = []
    try:
        s.relaxColdStart()
        es = s.electrostatics()
        while s.steps < TARGET:
            nxt = nextCycleTransition(s.steps, s._R.cycle_cfg)
            s.stepCycled(int(min(TARGET, nxt, (s.steps // 1000 + 1) * 1000) - s.steps))
            last = s.electrostatics()["last"]
            if s.steps % 1000 == 0 and last is not None:
                trace.append({"step": s.steps, "phase": s.cyclePhase, "box": float(s.box[0]), "N": s.N,
                              "alpha": last["alpha"]})
        nf = s.nonFiniteCount()
    finally:
        s.dispose()
    at = next(t for t in trace if t["step"] == TARGET)
    late = [t["alpha"] for t in trace if READ_FROM <= t["step"] <= TARGET]
    mean = float(np.mean(late))
    row = {"seed": seed, "I": ionic, "acidSoap": "off" if acid_soap == 0 else "shipped", "cutoff": es["cutoff"],
           "debyeLengths": es["debyeLengthsSpanned"], "alphaAt40000": at["alpha"], "pKaAppAt40000": apparentPKa(at["alpha"], PH),
           "phaseAt40000": at["phase"], "alphaMean36000to40000": mean, "pKaAppMean": apparentPKa(mean, PH),
           "alphaSdLate": float(np.std(late)), "nonFinite": nf, "seconds": time.time() - t0, "trace": trace}
    log(f"[es-live] seed={seed} I={ionic} acidSoap={row['acidSoap']} rc={row['cutoff']:.4f} ({row['debyeLengths']:.3f} lD) "
        f"phase@40000={at['phase']} alpha@40000={at['alpha']:.4f} pKa={row['pKaAppAt40000']:.3f} "
        f"mean36-40k={mean:.4f} pKa={row['pKaAppMean']:.3f} ({row['seconds']:.1f} s)")
    return row


def main():
    out_path = sys.argv[1] if len(sys.argv) > 1 else None
    seeds = [int(a) for a in sys.argv[2:]] or [19, 20, 21, 22, 23]
    vol = dry_box_volume()
    result = {"gate": "live salt shift of the apparent pKa, box 30, pH 6 (long-range-electrostatics-report.md:398-414, 495)",
              "command": "python tests/gates/es_salt_shift_live.py out.json " + " ".join(map(str, seeds)),
              "reference": {k: ({str(i): x for i, x in v.items()} if isinstance(v, dict) else v) for k, v in REFERENCE.items()},
              "runs": [], "summary": {}}
    for acid_soap in (0.0, None):
        for seed in seeds:
            for ionic in (0.01, 0.1):
                result["runs"].append(run_arm(seed, ionic, acid_soap, vol))
    for tag in ("off", "shipped"):
        rows = [r for r in result["runs"] if r["acidSoap"] == tag]
        shifts_at = [next(r for r in rows if r["seed"] == s and r["I"] == 0.01)["pKaAppAt40000"]
                     - next(r for r in rows if r["seed"] == s and r["I"] == 0.1)["pKaAppAt40000"] for s in seeds]
        shifts_mean = [next(r for r in rows if r["seed"] == s and r["I"] == 0.01)["pKaAppMean"]
                       - next(r for r in rows if r["seed"] == s and r["I"] == 0.1)["pKaAppMean"] for s in seeds]
        summ = {"seeds": seeds, "saltShiftAt40000": shifts_at, "saltShiftAt40000_mean": float(np.mean(shifts_at)),
                "saltShiftAt40000_sem": float(np.std(shifts_at, ddof=1) / np.sqrt(len(seeds))) if len(seeds) > 1 else None,
                "saltShiftLateMean": shifts_mean, "saltShiftLateMean_mean": float(np.mean(shifts_mean)),
                "saltShiftLateMean_sem": float(np.std(shifts_mean, ddof=1) / np.sqrt(len(seeds))) if len(seeds) > 1 else None}
        for ionic in (0.01, 0.1):
            a = [r["alphaAt40000"] for r in rows if r["I"] == ionic]
            summ[f"alphaAt40000_I{ionic}"] = {"mean": float(np.mean(a)), "sd": float(np.std(a, ddof=1)) if len(a) > 1 else None}
        result["summary"][tag] = summ
        log(f"[es-live] acidSoap={tag}: salt shift at 40000 {summ['saltShiftAt40000_mean']:.3f} "
            f"(+- {summ['saltShiftAt40000_sem']}), late mean {summ['saltShiftLateMean_mean']:.3f} "
            f"(+- {summ['saltShiftLateMean_sem']}); reference +0.709")
    text = json.dumps(result, indent=1)
    print(text)
    if out_path:
        with open(out_path, "w") as fh:
            fh.write(text)


if __name__ == "__main__":
    main()

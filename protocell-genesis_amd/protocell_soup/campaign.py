"""The soup campaign CLI (soup/cli/campaign.ts, soup/cli/campaign-config.ts) on this library:
run a soup to a step budget with periodic checkpoints in the reference's file layout, and continue
from the newest checkpoint of the same run configuration when one exists in --dir.

    python -m protocell_soup.campaign --label <name> --steps <n> --box <n> --start '{"C":300,"O":100}' [flags]

The flags, their defaults, the resume signature (configSignature), the newest-matching-checkpoint
scan and the crash-safe write are the reference's, so the two CLIs read and continue each other's
checkpoint directories. The run itself:
- create or resume (clay off, as the reference's CLI passes it);
- --relax on a fresh run only;
- --expandTo, a one-off growBoxTo with stepPhasesDEBUG probes before and after, a check that no
  particle moved, and a checkpoint;
- the step loop in blocks of --every (cut at each --minimiseAt step), each block followed by a
  checkpoint and a progress line (stage, aggregates, census, cycle phase, timings, the encapsulated
  water of the largest aggregate, and the charge statistics of a charged run).
SIGINT and SIGTERM stop the loop after the current block, which is checkpointed first.
Parcel confinement (--confineRadius, --confineStiffness) is not part of this build and is refused.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import signal
import sys
import time
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from .checkpoint import _js_number, decodeCheckpointResume, encodeCheckpoint, jsonStringify, writeCheckpointFile
from .rules import loadParams, loadSoup, wcaCutoff

GRID_DEBUG_N = 100  # steps each stepPhasesDEBUG group runs for the expansion probes (campaign.ts:37)


# ---- flags (campaign-config.ts:14-246) --------------------------------------------------------------
USAGE = """Usage:
  python -m protocell_soup.campaign --label <name> --steps <n> --box <n> --start '{"C":300,"O":100,"H":300,"M":20}' [flags]

Required:
  --label <name>    tells apart the checkpoint files of one directory (several campaigns may share --dir)
  --steps <n>       how many more steps to take in this call; on resume, counted from the step found, not from zero
  --box <n>         cubic box (one number, all three axes)
  --start <json>    starting composition by monomer id, for example '{"C":300,"O":100,"H":300,"M":20}'

Optional:
  --every <n>       steps between checkpoints (default from data/soup.json's checkpoint.everySteps)
  --dir <path>      checkpoint directory (default from data/soup.json's checkpoint.dir)
  --seed <n>        (default 1)
  --kT <n>          (default data/params.json thermostat.kT)
  --catalyst <n>    overrides the catalyst count separately from --start (like CreateSoupOpts.catalystCount)
  --cycle           enables dry-wet cycling (data/soup.json dryWetCycle) for this system
  --evaporate       the dry phase removes solvent beads from the system, rehydration returns them (requires --cycle)
  --relax           cold-start energy minimisation (data/soup.json coldStartRelax) before the first step;
                    only for a fresh run, silently skipped on resume
  --expandTo <n>            one-off ramped expansion of the live box to [n,n,n] right after creation/resume,
                            before the main step loop (idempotent: skipped if the box is already expanded)
  --expandRampSteps <n>     ramp steps for --expandTo (default 15)
  --expandRampRelaxSteps <n>  steps of ordinary dynamics between ramp increments (default 60)
  --cycles <n>      number of dry-wet cycles for this run (overrides data/soup.json
                    dryWetCycle.cycles; part of the resume signature: one cycle and six are different experiments)
  --minimiseAt <s1,s2,...>  global steps at which to apply energy minimisation mid-run
                    ("minimisation only" control, no cycling; part of the resume signature)
  --minimiseIterations <n>  iterations for each such minimisation (default coldStartRelax.iterations;
                    for the control, the number is taken from the [evaporation] lines of the cycled arm)
  --charge          enables screened electrostatics + head protonation equilibrium
                    (data/soup.json electrostatics) for this run; part of the resume signature
  --pH <n>          pH of the run (default data/soup.json electrostatics.pH); part of the signature
  --ionicStrength <n>  ionic strength, mol/L; sets the Debye length (default from the file);
                    part of the resume signature
  --confineRadius <n>, --confineStiffness <n>  parcel confinement: not part of this build (refused)"""


class _Parser(argparse.ArgumentParser):
    def error(self, message):
        raise ValueError(f"campaign: {message}")


def parseCliArgs(argv: Optional[Sequence[str]] = None) -> Dict[str, Any]:
    """parseCliArgs (campaign-config.ts:187-246) over argv (default sys.argv[1:])."""
    soup, params = loadSoup(), loadParams()
    if not soup.get("checkpoint"):
        raise ValueError("data/soup.json: missing section 'checkpoint' -- campaign.ts cannot choose the default interval/directory without it")
    ap = _Parser(add_help=False, allow_abbrev=False)
    for name in ("label", "steps", "every", "dir", "box", "seed", "kT", "start", "catalyst", "expandTo", "expandRampSteps",
                 "expandRampRelaxSteps", "cycles", "minimiseAt", "minimiseIterations", "pH", "ionicStrength",
                 "confineRadius", "confineStiffness"):
        ap.add_argument(f"--{name}", type=str)
    for name in ("cycle", "evaporate", "relax", "charge", "help"):
        ap.add_argument(f"--{name}", action="store_true")
    v = vars(ap.parse_args(list(sys.argv[1:] if argv is None else argv)))
    if v["help"]:
        return {"help": True}
    missing = [k for k in ("label", "steps", "box", "start") if v[k] is None]
    if missing:
        raise ValueError("campaign.ts: required flags missing: " + ", ".join(f"--{m}" for m in missing))
    num = float
    min_at: List[float] = []
    if v["minimiseAt"] is not None:
        min_at = sorted(num(x) for x in v["minimiseAt"].replace(",", " ").split() if x)
    return {
        "help": False,
        "label": v["label"],
        "steps": num(v["steps"]),
        "every": num(v["every"]) if v["every"] is not None else soup["checkpoint"]["everySteps"],
        "dir": v["dir"] if v["dir"] is not None else soup["checkpoint"]["dir"],
        "box": num(v["box"]),
        "seed": num(v["seed"]) if v["seed"] is not None else 1,
        "kT": num(v["kT"]) if v["kT"] is not None else params["thermostat"]["kT"],
        "start": json.loads(v["start"]),
        "catalyst": num(v["catalyst"]) if v["catalyst"] is not None else None,
        "cycle": v["cycle"],
        "evaporate": v["evaporate"],
        "relax": v["relax"],
        "expandTo": num(v["expandTo"]) if v["expandTo"] is not None else None,
        "expandRampSteps": num(v["expandRampSteps"]) if v["expandRampSteps"] is not None else 15,
        "expandRampRelaxSteps": num(v["expandRampRelaxSteps"]) if v["expandRampRelaxSteps"] is not None else 60,
        "cycles": num(v["cycles"]) if v["cycles"] is not None else None,
        "minimiseAt": min_at,
        "minimiseIterations": (num(v["minimiseIterations"]) if v["minimiseIterations"] is not None
                               else (soup.get("coldStartRelax") or {}).get("iterations", 200)),
        "charge": v["charge"],
        "pH": num(v["pH"]) if v["pH"] is not None else None,
        "ionicStrength": num(v["ionicStrength"]) if v["ionicStrength"] is not None else None,
        "confineRadius": num(v["confineRadius"]) if v["confineRadius"] is not None else None,
        "confineStiffness": num(v["confineStiffness"]) if v["confineStiffness"] is not None else 100,
    }


def _intish(x):
    """A JS number read from a flag: integral values as int, so they print and compare as JS does."""
    if isinstance(x, float) and x.is_integer():
        return int(x)
    return x


def runConfig(args: Dict[str, Any]) -> Dict[str, Any]:
    """The CheckpointConfig of a run (campaign.ts:55-80): keys in the reference's order, undefined
    ones left out as JSON.stringify leaves them out."""
    soup = loadSoup()
    b = _intish(args["box"])
    cfg: Dict[str, Any] = {"box": [b, b, b], "seed": _intish(args["seed"]), "kT": _intish(args["kT"]), "start": args["start"]}
    if args.get("catalyst") is not None:
        cfg["catalystCount"] = _intish(args["catalyst"])
    if args.get("cycle"):
        cfg["dryWetCycle"] = True
    if args.get("evaporate"):
        cfg["evaporateSolvent"] = True
    if args.get("cycles") is not None:
        cfg["dryWetCycles"] = _intish(args["cycles"])
    if args.get("minimiseAt"):
        cfg["minimiseAt"] = [_intish(x) for x in args["minimiseAt"]]
    if args.get("charge"):
        es = soup["electrostatics"]
        cfg["electrostatics"] = {"enabled": True,
                                 "pH": _intish(args["pH"]) if args.get("pH") is not None else es["pH"],
                                 "ionicStrengthMolar": (_intish(args["ionicStrength"]) if args.get("ionicStrength") is not None
                                                        else es["ionicStrengthMolar"])}
    if args.get("confineRadius") is not None:
        cfg["confine"] = {"radiusSigma": _intish(args["confineRadius"]), "stiffness": _intish(args["confineStiffness"])}
    return cfg


def _locale_key(s: str):
    """String.prototype.localeCompare order for the ASCII ids of data/soup.json: letters compared
    case-blind first, lowercase before uppercase on a tie."""
    return (s.lower(), [c.isupper() for c in s], s)


def configSignature(c: Dict[str, Any]) -> str:
    """configSignature (campaign-config.ts:252-283): the run identity a resume must match."""
    start = {k: (c.get("start") or {})[k] for k in sorted((c.get("start") or {}), key=_locale_key)}
    sig: Dict[str, Any] = {"box": c.get("box"), "seed": c.get("seed"), "kT": c.get("kT"), "start": start,
                           "catalystCount": c.get("catalystCount"), "dryWetCycle": c.get("dryWetCycle"),
                           "evaporateSolvent": c.get("evaporateSolvent")}
    if c.get("dryWetCycles") is not None:
        sig["dryWetCycles"] = c["dryWetCycles"]
    if c.get("minimiseAt"):
        sig["minimiseAt"] = c["minimiseAt"]
    if c.get("electrostatics") is not None:
        sig["electrostatics"] = c["electrostatics"]
    if c.get("confine") is not None:
        sig["confine"] = c["confine"]
    return jsonStringify(sig)


def findNewestMatchingCheckpoint(directory: str, label: str, sig: str) -> Optional[Dict[str, Any]]:
    """campaign-config.ts:285-309: among `<label>-step*.json` files whose config has this signature,
    the one with the highest globalStep; unreadable files are skipped."""
    if not os.path.isdir(directory):
        return None
    prefix = f"{label}-step"
    best = None
    for name in os.listdir(directory):
        if not name.startswith(prefix) or not name.endswith(".json"):
            continue
        path = os.path.join(directory, name)
        try:
            with open(path, encoding="utf-8") as fh:
                file = json.load(fh)
        except (OSError, ValueError):
            continue
        if configSignature(file.get("config") or {}) != sig:
            continue
        if best is None or file["globalStep"] > best["file"]["globalStep"]:
            best = {"path": path, "file": file}
    return best


# ---- output helpers ----------------------------------------------------------------------------------
def _fixed(x: float, d: int) -> str:
    return f"{x:.{d}f}"


def _exp(x: float, d: int) -> str:
    """Number.prototype.toExponential(d)."""
    if not math.isfinite(x):
        return "Infinity" if x > 0 else ("-Infinity" if x < 0 else "NaN")
    m, e = f"{x:.{d}e}".split("e")
    return f"{m}e{'+' if int(e) >= 0 else '-'}{abs(int(e))}"


def _fingerprint(p: np.ndarray) -> Dict[str, Any]:
    """campaign.ts:236-243: the running sum in float64 (element order) and every 3701st float."""
    a = np.asarray(p, dtype=np.float32).astype(np.float64)
    return {"sum": float(np.cumsum(a)[-1]) if a.size else 0.0, "sample": [float(x) for x in a[::3701]]}


def _create(cfg: Dict[str, Any], checkpoint: Optional[Dict[str, Any]]):
    from .sim import createSoup

    opts = {k: cfg[k] for k in ("box", "seed", "kT", "start", "catalystCount", "dryWetCycle", "dryWetCycles", "evaporateSolvent",
                                "electrostatics") if k in cfg}
    if checkpoint is not None:
        opts["resume"] = decodeCheckpointResume(checkpoint)
    opts["clay"] = False
    return createSoup(opts)


def progressLine(sys_, current: int, target: int, step_ms: int, ckpt_ms: int, saved: str, ctx=None) -> str:
    """The [campaign] progress line (campaign.ts:394-480) of a live system; progressMs is the time
    this function spends on the analysis."""
    from .analysis import detectStage
    from .electrostatics import apparentPKa, pairingStats

    t0 = time.time()
    stage, ev = detectStage(sys_, ctx)
    a = ev["aggregateAnalysis"]
    aggs = a["aggregates"]
    largest = None
    for g in aggs:  # the first of the largest, as the reference's reduce keeps it
        if largest is None or g["amphiphileCount"] > largest["amphiphileCount"]:
            largest = g
    inv = sys_.invariants()
    es_line = ""
    es = sys_.electrostatics()
    if es["enabled"]:
        p = loadParams()
        contact = wcaCutoff(p["sigma"] * p["beadSizes"]["head_head"])
        pr = pairingStats(sys_.particles(), sys_.charges(), sys_.box, sys_._R.es, contact)
        pka = apparentPKa(pr["alpha"], es["pH"])
        es_line = (f" rc_es={_fixed(es['cutoff'], 4)}(={_fixed(es['debyeLengthsSpanned'], 3)}lD, "
                   f"discarded={_fixed(es['discardedIntegratedFraction'], 4)}) pH={_js_number(es['pH'])} "
                   f"I={_js_number(es['ionicStrengthMolar'])} alpha={_fixed(pr['alpha'], 4)} "
                   f"pKaApp={'NaN' if math.isnan(pka) else _fixed(pka, 3)} paired={_fixed(pr['pairedFraction'], 4)} "
                   f"unlike={_fixed(pr['unlikeFraction'], 4)}(random {_fixed(pr['unlikeFractionRandom'], 4)}) sweeps={es['sweeps']}")
        if es.get("last"):
            es_line += (f" last(accepted={es['last']['accepted']}/{es['last']['attempts']} "
                        f"dEs={_fixed(es['last']['dEsMeanKT'], 4)}kT)")
    enc = largest.get("encapsulatedWater") if largest else None
    enc_line = (f" encH2O={enc['encapsulatedCount']}/{_fixed(float(enc['encapsulationThresholdCount']), 2)} "
                f"closed={'true' if enc['closed'] else 'false'} bulk_water_density={_fixed(enc['bulkWaterDensity'], 4)}"
                if enc else "")
    g = largest or {}
    shells = str(g["radialHeadShells"]) if largest else "n/a"
    progress_ms = int((time.time() - t0) * 1000)
    return (f"[campaign] step={current}/{target} stage={stage} aggregates={a['aggregateCount']} "
            f"largest={g.get('amphiphileCount', 0)} headShells={shells} cavityVolume={_fixed(g.get('cavityVolume', 0), 3)} "
            f"phase={sys_.cyclePhase}/{sys_.cycleIndex} box={_fixed(sys_.box[0], 4)} bonds={inv['bonds']} "
            f"census={jsonStringify(inv['monomers'])} flat={_fixed(g.get('flatnessRatio', 0), 4)} "
            f"inPl={_fixed(g.get('inPlaneSymmetry', 0), 4)} rg={_fixed(g.get('radiusOfGyration', 0), 3)} "
            f"histogram={jsonStringify(list(a['sizeHistogram'])[:6])} stepMs={step_ms} checkpointMs={ckpt_ms} "
            f"progressMs={progress_ms} saved={saved}" + enc_line + es_line)


def _probe(cfg, checkpoint, n, expand_to, ramp, relax_steps):
    """A throwaway system from the same start, optionally expanded, timed by stepPhasesDEBUG (campaign.ts:181-207)."""
    probe = _create(cfg, checkpoint)
    try:
        if expand_to is not None:
            probe.growBoxTo([expand_to] * 3, ramp, relax_steps)
        return {"debug": probe.stepPhasesDEBUG(n), "box": probe.box, "steps": probe.steps}
    finally:
        probe.dispose()


def _phase_ms(label: str, r: Dict[str, Any]) -> str:
    d = r["debug"]
    return (f"box={jsonStringify(r['box'])} steps={r['steps']} n={GRID_DEBUG_N} full={_fixed(d['full'], 4)}ms "
            f"gridBuild={_fixed(d['gridBuild'], 4)}ms force={_fixed(d['force'], 4)}ms "
            f"bondAttempts={_fixed(d['bondAttempts'], 4)}ms integration={_fixed(d['integration'], 4)}ms")


def run(args: Dict[str, Any], out=print) -> Dict[str, Any]:
    """main() of campaign.ts:47-520 for parsed flags; returns {step, stopped, saved}."""
    if args.get("confineRadius") is not None:
        raise ValueError("campaign: parcel confinement (--confineRadius) is not part of this build")
    cfg = runConfig(args)
    os.makedirs(args["dir"], exist_ok=True)
    sig = configSignature(cfg)
    found = findNewestMatchingCheckpoint(args["dir"], args["label"], sig)
    out(f"[campaign] resume label={args['label']} from {found['path']}, step={found['file']['globalStep']}" if found
        else f"[campaign] new run label={args['label']} (no matching checkpoints in {args['dir']})")
    state = {"stop": False}

    def on_signal(signum, _frame):
        out(f"[campaign] received {signal.Signals(signum).name} -- will save a checkpoint and stop after the current block")
        state["stop"] = True

    old = {s: signal.signal(s, on_signal) for s in (signal.SIGINT, signal.SIGTERM)}
    saved = None
    sys_ = None
    try:
        start_step = found["file"]["globalStep"] if found else 0
        target = int(start_step + args["steps"])
        sys_ = _create(cfg, found["file"] if found else None)
        out(f"[campaign] system ready N={sys_.N} start_step={sys_.steps} target={target}")
        if args.get("relax"):
            if start_step != 0:
                out(f"[campaign] --relax skipped: this is a resume from step={start_step}, minimisation is allowed only on a fresh start")
            else:
                r = sys_.relaxColdStart()
                out(f"[campaign] cold-start minimisation: iterations={r['iterations']} first_step={_js_number(r['maxDisplacementStart'])} "
                    f"total_displacement_bound={_fixed(r['displacementBound'], 4)} max|F| {_exp(r['maxForceBefore'], 4)} -> "
                    f"{_exp(r['maxForceAfter'], 4)} nonfinite_before={r['nonFiniteBefore']} nonfinite_after={r['nonFiniteAfter']} "
                    f"system_step={sys_.steps}")
        if args.get("expandTo") is not None:
            e = float(args["expandTo"])
            live = sys_.box
            if abs(live[0] - e) > 1e-6:
                ramp, rrelax = int(args["expandRampSteps"]), int(args["expandRampRelaxSteps"])
                out(f"[campaign] box expansion {jsonStringify(live)} -> [{_js_number(e)},{_js_number(e)},{_js_number(e)}] "
                    f"(growBoxTo, rampSteps={ramp}, rampRelaxSteps={rrelax})")
                ck = found["file"] if found else None
                before = _probe(cfg, ck, GRID_DEBUG_N, None, 0, 0)
                out("[campaign] before expansion (probe): " + _phase_ms("before", before))
                inv0, fp0 = sys_.invariants(), _fingerprint(sys_.particles())
                t0 = time.time()
                try:
                    sys_.growBoxTo([e, e, e], ramp, rrelax)
                except Exception as err:
                    out(f"[campaign] growBoxTo threw on an intermediate box: {err}")
                    out(f"[campaign] saving an emergency checkpoint at box={jsonStringify(sys_.box)}")
                    p = writeCheckpointFile(args["dir"], args["label"], encodeCheckpoint(sys_, cfg))
                    out(f"[campaign] emergency checkpoint saved: {p}")
                    raise
                expand_ms = int((time.time() - t0) * 1000)
                inv1, fp1 = sys_.invariants(), _fingerprint(sys_.particles())
                after = _probe(cfg, ck, GRID_DEBUG_N, e, ramp, rrelax)
                out(f"[campaign] after expansion ({expand_ms}ms wall clock, probe): " + _phase_ms("after", after))
                out(f"[campaign] the live system after expansion: box={jsonStringify(sys_.box)} steps={sys_.steps} "
                    f"invariantsBefore={jsonStringify(inv0)} invariantsAfter={jsonStringify(inv1)} "
                    f"fingerprintSumBefore={_js_number(fp0['sum'])} fingerprintSumAfter={_js_number(fp1['sum'])}")
                if inv0["monomers"] != inv1["monomers"] or inv0["charge"] != inv1["charge"]:
                    raise RuntimeError(f"[campaign] box expansion lost/added particles: {jsonStringify(inv0['monomers'])} -> "
                                       f"{jsonStringify(inv1['monomers'])}")
                if fp0 != fp1:
                    raise RuntimeError("[campaign] box expansion (growBoxTo) moved at least one particle -- the position "
                                       "fingerprint changed, and it must not have")
                out("[campaign] invariant confirmed: the per-monomer particle count, the charge and the position fingerprint "
                    "were not changed by the box expansion (growBoxTo)")
                p = writeCheckpointFile(args["dir"], args["label"], encodeCheckpoint(sys_, cfg))
                out(f"[campaign] checkpoint after expansion saved: {p}")
            else:
                out(f"[campaign] box already expanded to {jsonStringify(live)} -- skipping (idempotent resume)")
        current = sys_.steps
        due = sorted(x for x in args.get("minimiseAt", []) if current < x <= target)
        if args.get("minimiseAt"):
            out(f"[campaign] \"minimisation only\" control: steps={jsonStringify([_intish(x) for x in args['minimiseAt']])} "
                f"iterations_each={_js_number(args['minimiseIterations'])} still_due_in_this_call={jsonStringify([_intish(x) for x in due])}")
        from .analysis import AnalysisContext

        ctx = AnalysisContext()
        try:
            while current < target and not state["stop"]:
                nxt = next((x for x in due if x > current), None)
                chunk = int(min(args["every"], target - current, (nxt - current) if nxt is not None else math.inf))
                t0 = time.time()
                sys_.stepCycled(chunk)
                current = sys_.steps
                step_ms = int((time.time() - t0) * 1000)
                if nxt is not None and current == nxt:
                    m = sys_.minimiseNowDEBUG(int(args["minimiseIterations"]))
                    out(f"[campaign] mid-run minimisation at step={m['globalStep']} iterations={m['iterations']} "
                        f"max|F| {_exp(m['maxForceBefore'], 4)} -> {_exp(m['maxForceAfter'], 4)} "
                        f"displacement_bound={_fixed(m['displacementBound'], 4)}")
                t1 = time.time()
                file = encodeCheckpoint(sys_, cfg)
                ckpt_ms = int((time.time() - t1) * 1000)
                saved = writeCheckpointFile(args["dir"], args["label"], file)
                out(progressLine(sys_, current, target, step_ms, ckpt_ms, saved, ctx))
        finally:
            ctx.close()
        if state["stop"]:
            out(f"[campaign] stopped by signal at step={current} -- the next call with the same flags will continue from this point")
        else:
            out(f"[campaign] step budget fully completed: step={current}")
        return {"step": current, "stopped": state["stop"], "saved": saved}
    finally:
        for s, h in old.items():
            signal.signal(s, h)
        if sys_ is not None:
            sys_.dispose()


def main(argv: Optional[Sequence[str]] = None) -> int:
    try:
        args = parseCliArgs(argv)
        if args.get("help"):
            print(USAGE)
            return 0
        run(args, out=lambda s: print(s, flush=True))
        return 0
    except ValueError as err:
        if "required flags missing" in str(err):
            print(USAGE)
        print(err, file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())

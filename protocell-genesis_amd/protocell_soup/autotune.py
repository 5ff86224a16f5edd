"""Autotuner of the library's launch shapes.

Each tunable is a compile-time macro of csrc/soup_kernels.inl that changes how a kernel is laid out
on the wavefront, never what it computes (lists stay in the reference's order, MC decisions and RNG
streams are untouched, and the force sums stay in FP64 rounded once):

    FORCE_TPP    lanes per particle in the list force (4, 8 or 16)
    GROUP_BATCH  list entries per lane whose loads the list force issues together
    FAR_BATCH    far Debye-Hueckel chunks of 64 loaded together
    VL_BATCH     Verlet-build candidate chunks of 64 swept together
    ES_BATCH     DH-list-build candidate chunks of 64 loaded together
    VL_CELLS     the Verlet build as one workgroup per cell with the cell's neighbourhood staged in
                 LDS (1) or as one wavefront per particle gathering from HBM (0)
    VL_WAVES     wavefronts per workgroup of the per-cell Verlet build (8 or 16)

The reference's rebuild cadence is not tuned: the Verlet list feeds the bond Monte Carlo candidate
walk, so rebuildEvery is part of the model (data/soup.json verletList) and the MC decisions follow
it. The displacement-triggered schedule (verletOverride.rebuild = "drift"), where the lists follow
the drift instead, has one free parameter, how often the drift maximum is read back
(driftCheckEvery; its trigger follows from it, skin/2 less the displacement that many steps can
add). `schedule` times the campaign under each candidate and writes the fastest into the tuning
file's "rebuild" section, which the drift schedule takes as its default:

    python -m protocell_soup.autotune schedule [STEPS] [OUT]   # on the GPU

Two steps, so that compilation stays on a build host and timing on the GPU:

    python -m protocell_soup.autotune build DIR      # compile the variants (no GPU needed)
    python -m protocell_soup.autotune measure DIR    # time them on the GPU, write tuning.json

`build` compiles the current tuning and every variant that changes one macro from it. `measure`
times each variant's own kernel group (device events around the phase, as SoupSystem.profile
reads them) on the campaign soup (483,268 beads, 76 sigma, pH 7.0, 0.01 M), each variant in a
child process of its own, then keeps per macro the fastest value for its phase and writes
protocell_soup/tuning.json, which build_library compiles in.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import time
from typing import Dict, List

from .build import TUNING, build_library, tuning_defines

TUNABLES = {
    "FORCE_TPP": ((4, 8, 16), "force"),
    "GROUP_BATCH": ((2, 4, 8), "force"),
    "FAR_BATCH": ((8, 16, 32), "es_far"),
    "VL_BATCH": ((2, 4, 8), "verlet_build"),
    "ES_BATCH": ((4, 8, 16), "es_build"),
    "VL_CELLS": ((0, 1), "verlet_build"),
    "VL_WAVES": ((8, 16), "verlet_build"),
}
DEFAULTS = {"FORCE_TPP": 8, "GROUP_BATCH": 4, "FAR_BATCH": 16, "VL_BATCH": 4, "ES_BATCH": 8, "VL_CELLS": 1, "VL_WAVES": 16}
DRIFT_CHECK_CANDIDATES = (1, 2, 5, 10, 20)
CAMPAIGN = {"box": [76, 76, 76], "seed": 19, "kT": 1.1, "clay": False,
            "start": {"C": 55988, "O": 18659, "H": 55988, "M": 1452, "W": 351181},
            "electrostatics": {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01}}


def _name(d: Dict[str, int]) -> str:
    return "lib_" + "_".join(f"{k}{d[k]}" for k in sorted(d)) + ".so"


def variants(base: Dict[str, int]) -> List[Dict[str, int]]:
    """The base shapes and every shape that changes one macro from them."""
    out = [dict(base)]
    for k, (cands, _) in TUNABLES.items():
        for v in cands:
            if v != base[k]:
                out.append(dict(base, **{k: v}))
    return out


def current() -> Dict[str, int]:
    return dict(DEFAULTS, **tuning_defines())


def build(directory: str) -> List[str]:
    os.makedirs(directory, exist_ok=True)
    paths = []
    for d in variants(current()):
        path = os.path.join(directory, _name(d))
        build_library(out=path, defines=d)
        paths.append(path)
        print(f"[autotune] built {os.path.basename(path)}", flush=True)
    with open(os.path.join(directory, "variants.json"), "w") as fh:
        json.dump({"base": current(), "variants": variants(current())}, fh)
    return paths


def _time_one(steps: int) -> Dict[str, float]:
    """In a child process with SOUP_LIB_PATH set: the campaign soup, relaxed, `steps` profiled steps;
    device ms per launch of every phase."""
    import protocell_soup as ps

    s = ps.createSoup(CAMPAIGN)
    try:
        s.relaxColdStart()
        s.step(50)
        s.profile(True)
        s.step(steps)
        prof = s.profileRead()
        s.profile(False)
    finally:
        s.dispose()
    return {k: v["ms"] / v["count"] for k, v in prof.items() if v["count"]}


def measure(directory: str, steps: int = 200, out: str = TUNING) -> Dict[str, int]:
    with open(os.path.join(directory, "variants.json")) as fh:
        spec = json.load(fh)
    base = spec["base"]
    rows = []
    for d in spec["variants"]:
        path = os.path.join(directory, _name(d))
        pkg_root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        env = dict(os.environ, SOUP_LIB_PATH=path,
                   PYTHONPATH=os.pathsep.join(x for x in (pkg_root, os.environ.get("PYTHONPATH")) if x))
        t0 = time.time()
        res = subprocess.run([sys.executable, "-m", "protocell_soup.autotune", "time", str(steps)], env=env,
                             capture_output=True, text=True, timeout=300)
        if res.returncode != 0:
            raise RuntimeError(f"autotune: variant {d} failed:\n{res.stderr[-3000:]}")
        ms = json.loads(res.stdout.strip().splitlines()[-1])
        rows.append({"defines": d, "ms": ms})
        print(f"[autotune] {d} force={ms.get('force', 0):.4f} es_far={ms.get('es_far', 0):.4f} "
              f"verlet_build={ms.get('verlet_build', 0):.4f} es_build={ms.get('es_build', 0):.4f} ({time.time() - t0:.1f} s)",
              flush=True)
    best = dict(base)
    for k, (cands, phase) in TUNABLES.items():
        # the rows that differ from the base in k alone, and the base itself
        pool = [r for r in rows if all(r["defines"][j] == base[j] for j in base if j != k)]
        top = min(pool, key=lambda r: r["ms"].get(phase, float("inf")))
        best[k] = top["defines"][k]
    result = {"workload": "campaign soup 483,268 beads, 76 sigma, pH 7.0, 0.01 M", "steps": steps, "defines": best,
              "measurements": rows}
    if os.path.exists(out):  # keep the schedule's section
        with open(out) as fh:
            prev = json.load(fh)
        if "rebuild" in prev:
            result["rebuild"] = prev["rebuild"]
    with open(out, "w") as fh:
        json.dump(result, fh, indent=1)
    print(f"[autotune] chose {best} -> {out}", flush=True)
    return best


def _time_schedule(opts: Dict, steps: int) -> Dict[str, float]:
    """steps/s and rebuilds of the campaign under `opts` over `steps` steps (graphs on), after the
    cold-start relax and 200 steps."""
    import torch

    import protocell_soup as ps

    s = ps.createSoup(opts)
    try:
        s.relaxColdStart()
        s.step(200)
        r0 = s.rebuildCount()
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.step(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return {"steps_per_s": steps / dt, "rebuilds": s.rebuildCount() - r0}
    finally:
        s.dispose()


def schedule(steps: int = 1000, out: str = TUNING) -> int:
    """Times the campaign under the drift schedule at each driftCheckEvery candidate (and the
    reference's cadence, for the record) and writes the fastest candidate into `out`'s "rebuild"
    section, keeping its launch shapes."""
    rows = [dict(_time_schedule(CAMPAIGN, steps), schedule="cadence")]
    print(f"[autotune] cadence: {rows[0]['steps_per_s']:.1f} steps/s, {rows[0]['rebuilds']} rebuilds", flush=True)
    for e in DRIFT_CHECK_CANDIDATES:
        r = _time_schedule(dict(CAMPAIGN, verletOverride={"rebuild": "drift", "driftCheckEvery": e}), steps)
        rows.append(dict(r, schedule="drift", driftCheckEvery=e))
        print(f"[autotune] drift, check every {e}: {r['steps_per_s']:.1f} steps/s, {r['rebuilds']} rebuilds", flush=True)
    best = max((r for r in rows if r["schedule"] == "drift"), key=lambda r: r["steps_per_s"])
    data = {}
    if os.path.exists(out):
        with open(out) as fh:
            data = json.load(fh)
    data["rebuild"] = {"workload": "campaign soup 483,268 beads, 76 sigma, pH 7.0, 0.01 M", "steps": steps,
                       "driftCheckEvery": best["driftCheckEvery"], "measurements": rows}
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    print(f"[autotune] drift schedule: check every {best['driftCheckEvery']} -> {out}", flush=True)
    return best["driftCheckEvery"]


def tuned_drift_check_every(path: str = TUNING, default: int = 5) -> int:
    """The tuned driftCheckEvery of the drift schedule (`default` when none has been measured)."""
    if not os.path.exists(path):
        return default
    with open(path) as fh:
        v = json.load(fh).get("rebuild", {}).get("driftCheckEvery")
    return int(v) if isinstance(v, int) and v in DRIFT_CHECK_CANDIDATES else default


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in ("build", "measure", "time", "schedule"):
        print(__doc__)
        return 1
    if argv[0] == "build":
        build(argv[1])
    elif argv[0] == "schedule":
        schedule(int(argv[1]) if len(argv) > 1 else 1000, argv[2] if len(argv) > 2 else TUNING)
    elif argv[0] == "measure":
        measure(argv[1], int(argv[2]) if len(argv) > 2 else 200, argv[3] if len(argv) > 3 else TUNING)
    else:
        print(json.dumps(_time_one(int(argv[1]))))
    return 0


if __name__ == "__main__":
    sys.exit(main())

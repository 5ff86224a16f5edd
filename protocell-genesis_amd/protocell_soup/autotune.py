"""Autotuner of the library's launch shapes.

Each tunable is a compile-time macro of csrc/soup_kernels.inl that changes how a kernel is laid out
on the wavefront, never what it computes (lists stay in the reference's order, MC decisions and RNG
streams are untouched, and the force sums stay in FP64 rounded once):

    FORCE_TPP    lanes per particle in the list force (4, 8 or 16)
    GROUP_BATCH  list entries per lane whose loads the list force issues together
    FAR_BATCH    far Debye-Hueckel chunks of 64 loaded together
    VL_BATCH     Verlet-build candidate chunks of 64 loaded together
    ES_BATCH     DH-list-build candidate chunks of 64 loaded together

The rebuild cadence is not tuned: the Verlet list feeds the bond Monte Carlo candidate walk, so
rebuildEvery is part of the model (data/soup.json verletList) and the MC decisions follow it.

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
}
DEFAULTS = {"FORCE_TPP": 8, "GROUP_BATCH": 4, "FAR_BATCH": 16, "VL_BATCH": 4, "ES_BATCH": 8}
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
    with open(out, "w") as fh:
        json.dump(result, fh, indent=1)
    print(f"[autotune] chose {best} -> {out}", flush=True)
    return best


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in ("build", "measure", "time"):
        print(__doc__)
        return 1
    if argv[0] == "build":
        build(argv[1])
    elif argv[0] == "measure":
        measure(argv[1], int(argv[2]) if len(argv) > 2 else 200, argv[3] if len(argv) > 3 else TUNING)
    else:
        print(json.dumps(_time_one(int(argv[1]))))
    return 0


if __name__ == "__main__":
    sys.exit(main())

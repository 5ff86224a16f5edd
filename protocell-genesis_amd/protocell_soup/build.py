"""Builds libsoup_hip.so in-tree for gfx950 with hipcc (no JIT cache, no setup.py)."""
from __future__ import annotations

import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB = os.path.join(LIB_DIR, "libsoup_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # exact f32 op order in the parity kernels; the force kernels opt back in locally
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# launch shapes chosen by protocell_soup.autotune for the campaign workload (macro -> value)
TUNING = os.path.join(PKG_ROOT, "protocell_soup", "tuning.json")


def tuning_defines(path: str = TUNING) -> dict:
    """The tuned launch-shape macros, {} when no tuning file is present."""
    if not os.path.exists(path):
        return {}
    import json

    with open(path, encoding="utf-8") as fh:
        return {k: int(v) for k, v in json.load(fh).get("defines", {}).items()}


def build_library(verbose: bool = False, out: str = LIB, defines: dict = None) -> str:
    """hipcc of csrc/soup_lib.hip to `out`, with the tuned launch shapes (or `defines` instead)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        hipcc = "hipcc"
    tmp = out + ".tmp"
    d = tuning_defines() if defines is None else defines
    cmd = [hipcc, *HIPCC_FLAGS, *[f"-D{k}={v}u" for k, v in sorted(d.items())], "-I", os.path.join(PKG_ROOT, "..", "include"),
           "-o", tmp, os.path.join(CSRC, "soup_lib.hip")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-6000:])
    if verbose and res.stderr:
        print(res.stderr[-2000:])
    os.replace(tmp, out)
    return out


NAPI_SRC = os.path.join(PKG_ROOT, "napi", "soup_napi.cc")
NAPI_OUT = os.path.join(LIB_DIR, "soup_napi.node")
NODE_INCLUDE = "/usr/include/node"


def build_napi(out: str = NAPI_OUT) -> str:
    """g++ of napi/soup_napi.cc against Node's node_api.h, linked to libsoup_hip.so beside it (rpath
    $ORIGIN). Returns the path, or "" where Node's headers are not installed."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return ""
    cxx = os.environ.get("CXX", "g++")
    tmp = out + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", NODE_INCLUDE, "-I",
           os.path.join(PKG_ROOT, "..", "include"), NAPI_SRC, "-L", LIB_DIR, "-lsoup_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ of the N-API addon failed:\n" + res.stderr[-4000:])
    os.replace(tmp, out)
    return out

"""A resolved soup as files, for a host outside Python (the Node host in napi/soup.js).

createSoup's resolution (soup/src/sim.ts:96-278, restated in sim.py) produces the SoupConfig the
library takes and the initial state. writeBundle stores both in a directory:
- config.bin: the SoupConfig bytes (the C layout of include/soup_abi.h);
- one raw little-endian file per initial-state array (positions.f32, bondSlots.u32, ...);
- meta.json: the particle count, the rule count, the scalars of SoupInitState, the cold-start relax
  defaults and the array files with their element types.
The reference's TypeScript host would pack the same bytes from its own resolution; the bundle is the
handover for callers that resolve here and step elsewhere.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Any, Dict, Union

import numpy as np

from .sim import Resolved, resolve

# array name -> (file suffix, dtype)
ARRAYS = {
    "positions": ("f32", np.float32), "velocities": ("f32", np.float32), "bondSlots": ("u32", np.uint32),
    "centerLink": ("u32", np.uint32), "centerHeldSteps": ("u32", np.uint32), "frozen": ("u32", np.uint32),
    "bondRng": ("u32", np.uint32), "thermoRng": ("u32", np.uint32), "charges": ("f32", np.float32),
    "events": ("u32", np.uint32), "desorbEvents": ("u32", np.uint32),
}


def writeBundle(opts_or_resolved: Union[Dict[str, Any], Resolved], directory: str) -> Dict[str, Any]:
    """Resolves createSoup options (or takes a Resolved) and writes the bundle; returns meta."""
    R = opts_or_resolved if isinstance(opts_or_resolved, Resolved) else resolve(opts_or_resolved)
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "config.bin"), "wb") as fh:
        fh.write(C.string_at(C.addressof(R.cfg), C.sizeof(R.cfg)))
    arrays = dict(R.init)
    arrays["charges"] = R.charges0
    arrays["events"] = R.events0
    files = {}
    for name, (suffix, dt) in ARRAYS.items():
        a = np.ascontiguousarray(arrays[name], dtype=dt).reshape(-1)
        fn = f"{name}.{suffix}"
        a.astype(a.dtype.newbyteorder("<"), copy=False).tofile(os.path.join(directory, fn))
        files[name] = {"file": fn, "type": suffix, "length": int(a.size)}
    relax = R.soup.get("coldStartRelax") or {}
    meta = {
        "N": int(R.N), "nRules": int(R.cfg.n_rules), "configBytes": C.sizeof(R.cfg),
        "globalStep": int(R.initial_step),
        "nextSweepAt": -1 if R.next_sweep_at >= 2 ** 64 - 1 else int(R.next_sweep_at),
        "protonationRng": int(R.prng) & 0xFFFFFFFF,
        "sigma": float(R.p["sigma"]),
        "coldStartRelax": {"iterations": relax.get("iterations"), "maxDisplacementSigma": relax.get("maxDisplacementSigma")},
        "arrays": files,
    }
    with open(os.path.join(directory, "meta.json"), "w", encoding="utf-8") as fh:
        json.dump(meta, fh, indent=1)
    return meta

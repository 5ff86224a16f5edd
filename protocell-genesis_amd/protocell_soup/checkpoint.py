"""Checkpoint/resume in the reference's file layout (soup/src/checkpoint.ts:110-237).

``encodeCheckpoint(sys, config)`` turns a live system's mutable state into the reference's
``CheckpointFile`` object (version 1, base64 of the raw little-endian typed-array bytes, keys in
the reference's order); ``decodeCheckpointResume(file)`` gives back ``createSoup``'s ``resume``
shape, so a resume is ``createSoup({**file["config"], "resume": decodeCheckpointResume(file)})``.
``jsonStringify`` writes the text JSON.stringify writes (no whitespace, JavaScript number
formatting), so files written here and by soup/cli/campaign-config.ts:317-323 are interchangeable
byte for byte apart from ``createdAt``.
"""
from __future__ import annotations

import base64
import datetime
import json
import math
import os
from typing import Any, Dict

import numpy as np


# ---- typed arrays <-> base64 (checkpoint.ts:37-63) ---------------------------------------------
def encodeTyped(arr: np.ndarray) -> str:
    """Base64 of the array's bytes in little-endian order (a JS typed array's own byte order)."""
    a = np.ascontiguousarray(arr)
    if a.dtype.byteorder == ">":
        a = a.byteswap().view(a.dtype.newbyteorder("<"))
    return base64.b64encode(a.tobytes()).decode("ascii")


def decodeFloat32(b64: str) -> np.ndarray:
    return np.frombuffer(base64.b64decode(b64), dtype="<f4").copy()


def decodeUint32(b64: str) -> np.ndarray:
    return np.frombuffer(base64.b64decode(b64), dtype="<u4").copy()


# ---- JSON.stringify ------------------------------------------------------------------------------
def _js_number(x) -> str:
    """Number::toString (ECMA-262 6.1.6.1.20) from the shortest round-trip digits."""
    if isinstance(x, bool):
        return "true" if x else "false"
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if not math.isfinite(x):
        return "null"  # JSON.stringify writes NaN and +-Infinity as null
    if x == 0.0:
        return "0"
    sign = "-" if x < 0 else ""
    mant, _, exp = repr(abs(x)).partition("e")
    e10 = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    # value = 0.digits * 10^n
    n = len(ip.lstrip("0")) + e10 if ip.strip("0") else e10 - (len(fp) - len(fp.lstrip("0")))
    digits = digits.rstrip("0") or "0"
    k = len(digits)
    if k <= n <= 21:
        s = digits + "0" * (n - k)
    elif 0 < n <= 21:
        s = digits[:n] + "." + digits[n:]
    elif -6 < n <= 0:
        s = "0." + "0" * (-n) + digits
    else:
        e = n - 1
        s = digits[0] + ("." + digits[1:] if k > 1 else "") + "e" + ("+" if e >= 0 else "-") + str(abs(e))
    return sign + s


def jsonStringify(obj: Any) -> str:
    """JSON.stringify(obj) for plain data: dicts keep their insertion order, no whitespace."""
    if obj is None:
        return "null"
    if isinstance(obj, (bool, np.bool_)):
        return "true" if obj else "false"
    if isinstance(obj, (int, float, np.integer, np.floating)):
        return _js_number(obj)
    if isinstance(obj, str):
        return json.dumps(obj, ensure_ascii=False)
    if isinstance(obj, dict):
        return "{" + ",".join(json.dumps(str(k), ensure_ascii=False) + ":" + jsonStringify(v)
                              for k, v in obj.items() if v is not _UNDEFINED) + "}"
    if isinstance(obj, (list, tuple, np.ndarray)):
        return "[" + ",".join(jsonStringify(v) for v in obj) + "]"
    raise TypeError(f"jsonStringify: unsupported value of type {type(obj).__name__}")


_UNDEFINED = object()  # a key JSON.stringify omits (an undefined property)


# ---- encode / decode (checkpoint.ts:181-237) ----------------------------------------------------
def _iso_now() -> str:
    t = datetime.datetime.now(datetime.timezone.utc)
    return t.strftime("%Y-%m-%dT%H:%M:%S.") + f"{t.microsecond // 1000:03d}Z"


def encodeCheckpoint(sys, config: Dict[str, Any]) -> Dict[str, Any]:
    """encodeCheckpoint (checkpoint.ts:181-217): pure readbacks, no device work that changes the
    trajectory. ``sys`` needs the CheckpointableSystem accessors (checkpoint.ts:158-172)."""
    positions = sys.particles()
    rng = sys.rngState()
    return {
        "version": 1,
        "createdAt": _iso_now(),
        "config": config,
        "N": int(positions.size // 4),
        "globalStep": int(sys.steps),
        "liveBox": [float(v) for v in sys.box],
        "cyclePhase": sys.cyclePhase,
        "cycleIndex": int(sys.cycleIndex),
        "activeCounts": sys.invariants()["monomers"],
        "events": sys.events(),
        "desorbEvents": sys.desorbEvents(),
        "positionsB64": encodeTyped(positions.astype(np.float32)),
        "velocitiesB64": encodeTyped(sys.velocities().astype(np.float32)),
        "bondSlotsB64": encodeTyped(sys.bondSlots().astype(np.uint32)),
        "centerLinkB64": encodeTyped(sys.centerLinks().astype(np.uint32)),
        "centerHeldStepsB64": encodeTyped(sys.centerHeldSteps().astype(np.uint32)),
        "bondRngB64": encodeTyped(np.asarray(rng["bond"], dtype=np.uint32)),
        "thermoRngB64": encodeTyped(np.asarray(rng["thermo"], dtype=np.uint32)),
        "chargesB64": encodeTyped(sys.charges().astype(np.float32)),
        "protonationRng": int(sys.protonationRngState()),
    }


def decodeCheckpointResume(file: Dict[str, Any]) -> Dict[str, Any]:
    """decodeCheckpointResume (checkpoint.ts:222-237): createSoup's ``resume``."""
    if file.get("version") != 1:
        raise ValueError(f"checkpoint: unsupported version {file.get('version')!r} (expected 1)")
    return {
        "globalStep": file["globalStep"],
        "liveBox": file["liveBox"],
        "activeCounts": file.get("activeCounts"),
        "positions": decodeFloat32(file["positionsB64"]),
        "velocities": decodeFloat32(file["velocitiesB64"]),
        "bondSlots": decodeUint32(file["bondSlotsB64"]),
        "centerLink": decodeUint32(file["centerLinkB64"]),
        "centerHeldSteps": decodeUint32(file["centerHeldStepsB64"]),
        "desorbEvents": np.array([file["desorbEvents"]["stretch"], file["desorbEvents"]["timeout"]], dtype=np.uint32),
        "bondRng": decodeUint32(file["bondRngB64"]),
        "thermoRng": decodeUint32(file["thermoRngB64"]),
        "events": file["events"],
        "charges": decodeFloat32(file["chargesB64"]) if file.get("chargesB64") is not None else None,
        "protonationRng": file.get("protonationRng"),
    }


# ---- files (soup/cli/campaign-config.ts:317-323) -------------------------------------------------
def writeCheckpointFile(directory: str, label: str, file: Dict[str, Any]) -> str:
    """``<dir>/<label>-step<globalStep>.json``, written to a temporary name and renamed (atomic on
    one filesystem), so a half-written file is never taken for a checkpoint."""
    final = os.path.join(directory, f"{label}-step{file['globalStep']}.json")
    tmp = final + ".tmp"
    with open(tmp, "w", encoding="utf-8") as fh:
        fh.write(jsonStringify(file))
    os.replace(tmp, final)
    return final


def readCheckpointFile(path: str) -> Dict[str, Any]:
    with open(path, encoding="utf-8") as fh:
        return json.load(fh)

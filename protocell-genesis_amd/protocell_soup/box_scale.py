"""Box changes of a live soup: applyBoxScaleOnce, the ramped scaleBoxTo, growBoxTo and the dry-wet
cycle that drives them (soup/src/soup-box-scale.ts, soup/src/soup-box-scale-math.ts).

Every box change keeps the reference's order of work:
- the rigid-molecule map of the current configuration (scaleMoleculesRigid, native host code),
  optionally with proximity edges making a cohesive cluster one rigid unit;
- the bonded-distance self-check (every rigid edge keeps its length to 1e-3 sigma);
- the new grid from planSoupGrid;
- soup_set_box on the device: positions, grid reallocation when it grew, list rebuild, force.
growBoxTo moves nothing: it adopts each box of its ramp and rebuilds the lists, without a force
pass, as the reference's does.

The dry-wet cycle is resolved at creation (deriveCycleConfig): the wet box is the creation box, the
dry box holds the creation census at dryWetCycle.targetDryDensity. stepCycled steps to each
transition of the schedule and ramps the box there over dryWetCycle.rampSteps log-linear
increments, each followed by rampRelaxSteps of ordinary steps.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np

from . import native
from .plan import planSoupGrid

Box = Tuple[float, float, float]


# ---- schedule (soup-box-scale-math.ts:251-290) -----------------------------------------------------
@dataclass
class CycleSchedule:
    periodSteps: int
    dryFraction: float
    cycles: int


def cycleTransitionSteps(cfg: CycleSchedule):
    """Wet -> dry at k * period + wet length, dry -> wet at (k + 1) * period, for each cycle k."""
    wet_len = cfg.periodSteps * (1 - cfg.dryFraction)
    out = []
    for k in range(cfg.cycles):
        out.append(k * cfg.periodSteps + wet_len)
        out.append((k + 1) * cfg.periodSteps)
    return out


def cyclePhaseAt(step: int, cfg: CycleSchedule) -> Dict[str, Any]:
    """The phase and 1-based cycle of a step count; cycleIndex 0 once every cycle has run."""
    total = cfg.periodSteps * cfg.cycles
    if step >= total:
        return {"phase": "wet", "cycleIndex": 0}
    within = step % cfg.periodSteps
    wet_len = cfg.periodSteps * (1 - cfg.dryFraction)
    return {"phase": "wet" if within < wet_len else "dry", "cycleIndex": math.floor(step / cfg.periodSteps) + 1}


def nextCycleTransition(step: int, cfg: CycleSchedule) -> float:
    """The smallest transition step strictly after `step` (math.inf once cycling is over)."""
    best = math.inf
    for t in cycleTransitionSteps(cfg):
        if step < t < best:
            best = t
    return best


def computeDryBox(box: Sequence[float], N: int, targetDryDensity: float) -> Box:
    """soup-box-scale-math.ts:226-242: isotropic contraction to N / targetDryDensity."""
    wet_volume = box[0] * box[1] * box[2]
    dry_volume = N / targetDryDensity
    scale = np.cbrt(dry_volume / wet_volume)
    return (box[0] * float(scale), box[1] * float(scale), box[2] * float(scale))


def deriveCycleConfig(soup: Dict[str, Any], opts: Dict[str, Any], box: Sequence[float], N: int,
                      startCounts: Dict[str, int], dryBoxOverride: Optional[Sequence[float]] = None
                      ) -> Tuple[Optional[CycleSchedule], Box]:
    """soup-box-scale-math.ts:316-361 for unconfined systems; ``dryBoxOverride`` is the dry box of an
    evaporating cycle (sized for what is left, planEvaporation), None the box for the census N."""
    enabled = opts.get("dryWetCycle", soup["dryWetCycle"]["enabled"])
    box = tuple(float(x) for x in box)
    if not enabled:
        return None, box
    dwc = soup["dryWetCycle"]
    wet_density = N / (box[0] * box[1] * box[2])
    if dwc["targetDryDensity"] <= wet_density:
        raise ValueError(
            f"data/soup.json: dryWetCycle.targetDryDensity={_js_num(dwc['targetDryDensity'])} does not exceed the current broth "
            f"density {wet_density:.4f} (N={N}, box=[{','.join(_js_num(x) for x in box)}]) -- the dry phase must concentrate, not dilute")
    dry_box = tuple(float(x) for x in dryBoxOverride) if dryBoxOverride is not None else computeDryBox(box, N, dwc["targetDryDensity"])
    dry_plan = planSoupGrid(dry_box, startCounts)
    if not dry_plan.valid:
        raise ValueError(dry_plan.reason)
    cycles = opts.get("dryWetCycles", dwc["cycles"])
    if not isinstance(cycles, int) or isinstance(cycles, bool) or cycles < 1:
        raise ValueError(f"the number of dry-wet cycles must be an integer >= 1, got {cycles}")
    return CycleSchedule(dwc["periodSteps"], dwc["dryFraction"], cycles), dry_box


def _js_num(x: float) -> str:
    return str(int(x)) if float(x).is_integer() else repr(float(x))


# ---- the box change itself -----------------------------------------------------------------------
def _js_round(x: np.ndarray) -> np.ndarray:
    r = np.floor(x)
    return r + (x - r >= 0.5)


def mi3Distance(p: np.ndarray, box: Sequence[float], i: np.ndarray, j: np.ndarray) -> np.ndarray:
    """soup-box-scale-math.ts:296-301 for arrays of pairs: FP64 from the f32 coordinates."""
    p = p.reshape(-1, 4)
    b = np.asarray(box, np.float64)
    d = p[i, :3].astype(np.float64) - p[j, :3].astype(np.float64)
    d = d - _js_round(d / b) * b
    return np.sqrt((d * d).sum(axis=1))


def proximityPairs(positions: np.ndarray, box: Sequence[float], cutoff: float, eligible: np.ndarray) -> np.ndarray:
    """soup-box-scale-math.ts:162-215 through the library's host implementation."""
    lib = native.load_library()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    n = pos.size // 4
    el = np.ascontiguousarray(eligible, dtype=np.uint8)
    if el.size != n:
        raise ValueError("proximityPairs: eligible must hold one flag per particle")
    bx = (C.c_double * 3)(*[float(x) for x in box])
    count = C.c_uint64()
    r = lib.soup_proximity_pairs(pos.ctypes.data_as(C.c_void_p), C.c_uint32(n), bx, C.c_double(float(cutoff)),
                                 el.ctypes.data_as(C.c_void_p), None, C.c_uint64(0), C.byref(count))
    if r != 0:
        raise ValueError("proximityPairs: invalid box or cutoff")
    out = np.empty(2 * int(count.value), dtype=np.uint32)
    if out.size:
        lib.soup_proximity_pairs(pos.ctypes.data_as(C.c_void_p), C.c_uint32(n), bx, C.c_double(float(cutoff)),
                                 el.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_uint64(int(count.value)),
                                 C.byref(count))
    return out


def _box_str(b: Sequence[float]) -> str:
    return ",".join(_js_num(x) for x in b)


def _adopt_box(sys_, target: Box, positions: Optional[np.ndarray], flags: int) -> None:
    R = sys_._R
    plan = planSoupGrid(target, R.start)
    if not plan.valid:
        raise ValueError(plan.reason)
    geo = native.SoupConfig()
    C.memmove(C.byref(geo), C.byref(R.cfg), C.sizeof(native.SoupConfig))
    for a in range(3):
        geo.box[a] = target[a]
        geo.dims[a] = plan.dims[a]
    geo.walk_radius = plan.effectiveWalkRadius
    ptr = positions.ctypes.data_as(C.c_void_p) if positions is not None else None
    sys_._call(sys_._lib.soup_set_box, C.byref(geo), ptr, C.c_uint32(flags))
    R.cfg = geo
    R.live_box = target
    R.plan = plan


def applyBoxScaleOnce(sys_, targetBox: Sequence[float], cohesion: Optional[Dict[str, Any]] = None) -> None:
    """soup-box-scale.ts:81-201: the rigid map of the current configuration into `targetBox` (bonded
    molecules, plus proximity-connected clusters when `cohesion` = {cutoff, memberIndices?} is given),
    the self-check of every rigid edge, then the new grid, the rebuild and the force on the device."""
    R = sys_._R
    frozen = int((sys_.frozen() != 0).sum())
    if frozen > 0:
        raise ValueError(
            f"changing the box is impossible on a system with a mineral sheet: {frozen} immobile beads "
            "(areaMove/scaleBoxTo/dryWetCycle require CreateSoupOpts.clay=false, see data/soup.json clay.basis §7)")
    from .area_move import scaleMoleculesRigid

    target = tuple(float(x) for x in targetBox)
    old = tuple(float(x) for x in R.live_box)
    before = sys_.particles()
    edges = sys_.bonds()
    if cohesion is not None:
        eligible = np.zeros(R.N, dtype=np.uint8)
        members = cohesion.get("memberIndices")
        if members is not None:
            eligible[np.asarray(members, dtype=np.int64)] = 1
        else:
            eligible[:] = 1
        prox = proximityPairs(before, old, float(cohesion["cutoff"]), eligible)
        edges = np.concatenate([edges, prox]).astype(np.uint32)
    after = scaleMoleculesRigid(before, edges, old, target)
    if edges.size:
        i, j = edges[0::2].astype(np.int64), edges[1::2].astype(np.int64)
        d0 = mi3Distance(before, old, i, j)
        d1 = mi3Distance(after, target, i, j)
        bad = np.nonzero(np.abs(d1 - d0) > 1e-3)[0]
        if bad.size:
            k = int(bad[0])
            raise RuntimeError(
                f"applyBoxScale: an intramolecular distance changed ({d0[k]:.6f} -> {d1[k]:.6f}) for bond {i[k]}-{j[k]} "
                f"at box [{_box_str(old)}] -> [{_box_str(target)}] -- scaleMoleculesRigid is broken")
    _adopt_box(sys_, target, after, native.SET_BOX_FORCE)


def _ramp(fromBox: Box, targetBox: Box, s: int, rampSteps: int) -> Box:
    """Increment s of a log-linear ramp; the last one lands on the target exactly."""
    if s == rampSteps:
        return targetBox
    return tuple(fromBox[a] * math.exp((math.log(targetBox[a] / fromBox[a]) * s) / rampSteps) for a in range(3))


def applyBoxScale(sys_, targetBox: Sequence[float], rampSteps: int, rampRelaxSteps: int,
                  cohesion: Optional[Dict[str, Any]] = None) -> None:
    """soup-box-scale.ts:218-246: rampSteps increments of applyBoxScaleOnce, each but the last
    followed by rampRelaxSteps ordinary steps."""
    from_box = tuple(float(x) for x in sys_._R.live_box)
    target = tuple(float(x) for x in targetBox)
    for s in range(1, rampSteps + 1):
        applyBoxScaleOnce(sys_, _ramp(from_box, target, s, rampSteps), cohesion)
        if s < rampSteps and rampRelaxSteps > 0:
            sys_.step(rampRelaxSteps)


def growBoxTo(sys_, targetBox: Sequence[float], rampSteps: int, rampRelaxSteps: int) -> None:
    """soup-box-scale.ts:318-373: every increment of the ramp adopts its box and rebuilds the lists
    without moving any particle (and without a force pass); each but the last is followed by
    rampRelaxSteps ordinary steps."""
    from_box = tuple(float(x) for x in sys_._R.live_box)
    target = tuple(float(x) for x in targetBox)
    for s in range(1, rampSteps + 1):
        _adopt_box(sys_, _ramp(from_box, target, s, rampSteps), None, 0)
        if s < rampSteps and rampRelaxSteps > 0:
            sys_.step(rampRelaxSteps)


def stepCycled(sys_, n: int) -> None:
    """makeStepCycled (soup-box-scale.ts:478-532): n steps through the dry-wet schedule; at each
    transition the box ramps to the dry box or back to the wet (creation) box, and on an evaporating
    system the solvent leaves or returns on the way (protocell_soup.evaporate)."""
    R = sys_._R
    cfg = R.cycle_cfg
    if cfg is None:
        sys_.step(n)
        return
    dwc = R.soup["dryWetCycle"]
    remaining = int(n)
    while remaining > 0:
        nxt = nextCycleTransition(sys_.steps, cfg)
        advance = remaining if nxt == math.inf else min(remaining, nxt - sys_.steps)
        if advance > 0:
            sys_.step(int(advance))
            remaining -= int(advance)
        if nxt != math.inf and sys_.steps == nxt:
            ph = cyclePhaseAt(sys_.steps, cfg)
            if getattr(R, "evap", None) is not None:
                # the solvent leaves and returns at the same transition points
                from .evaporate import applyEvaporatingTransition

                rungs: list = []
                reports = applyEvaporatingTransition(sys_, R.evap, ph["phase"], log=rungs.append)
                R.last_transition = {"phase": ph["phase"], "step": int(nxt), "rungs": rungs, "rehydration": reports}
            else:
                applyBoxScale(sys_, R.dry_box if ph["phase"] == "dry" else R.box, dwc["rampSteps"], dwc["rampRelaxSteps"])
            R.cycle_phase = ph["phase"]
            R.cycle_index = ph["cycleIndex"]
        elif advance == 0 and remaining > 0:
            raise RuntimeError("stepCycled: the cycle schedule looped -- nextCycleTransition returned the current globalStep")

"""Solvent evaporation and rehydration on a live soup (soup/src/soup-evaporate-plan.ts,
soup/src/soup-evaporate.ts, and the evaporating transition of soup/src/soup-box-scale.ts:356-465).

The solvent block is the last non-empty block of the composition, so removing solvent shrinks the
live count from the top and no other particle index moves: the bond graph, the valence slots, the
catalyst links and every organic bead are untouched by construction. The device buffers keep the
creation census (SoupConfig.capacity), so rehydration writes the returning beads back into the same
index range.

On the device a change of the live count is one call, soup_reset_particles: the host reads the state
in particle order, keeps its first n particles (evaporation) or appends the returning beads
(rehydration), and the library re-uploads it with the identity slot order, rebuilds the lists,
recomputes the force and runs the finiteness and list guards -- setActiveCount's post-conditions
(soup-evaporate.ts:58-80). The step counter, the event and desorption counters and the constant-pH
stream carry over.

The insertion sampler (best of 64 uniform candidates by distance to everything already there) is
host code in the library, soup_sample_insertion, in FP64 as the reference computes it; its
mulberry32 stream continues into the velocity draws of the returning beads.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import native
from .init_state import NONE_U32, Mulberry32, _bm
from .rules import wcaCutoff

Box = Tuple[float, float, float]


@dataclass
class EvaporationPlan:
    """planEvaporation's result (soup-evaporate-plan.ts:27-46)."""

    solventKind: int
    solventBlockStart: int
    wetSolventCount: int
    drySolventCount: int
    organicCount: int
    wetBox: Box
    dryBox: Box
    rampSteps: int
    minSeparationSigma: float
    relaxIterations: int
    concentrationFactor: float


def _js_round_scalar(x: float) -> int:
    r = math.floor(x)
    return int(r + 1) if x - r >= 0.5 else int(r)


def planEvaporation(soup: Dict[str, Any], p: Dict[str, Any], box: Sequence[float], startCounts: Dict[str, int]) -> EvaporationPlan:
    """soup-evaporate-plan.ts:51-121 for unconfined systems: the solvent block, the dry census, the
    dry box for what is left, and the minimiser budget an inserted bead needs."""
    from .box_scale import computeDryBox

    dwc = soup["dryWetCycle"]
    if dwc.get("residualSolventFraction") is None or dwc.get("evaporationRampSteps") is None \
            or dwc.get("insertionMinSeparationSigma") is None:
        raise ValueError("data/soup.json: solvent evaporation requested, but dryWetCycle does not contain its fields "
                         "(residualSolventFraction / evaporationRampSteps / insertionMinSeparationSigma)")
    water = soup["solvent"]["waterId"]
    ids = [m["id"] for m in soup["monomers"]]
    if water not in ids:
        raise ValueError(f"data/soup.json: solvent.waterId={water} not found among monomers")
    solvent_kind = ids.index(water)
    counts = [int(startCounts.get(m, 0) or 0) for m in ids]
    for k in range(solvent_kind + 1, len(counts)):
        if counts[k] > 0:
            raise ValueError(
                f"solvent evaporation: the solvent block ({water}, kind {solvent_kind}) must be the last non-empty block of the "
                f"composition, but kind {ids[k]} has {counts[k]} particles -- truncation would shift the indices of non-water "
                "particles and break the bond graph")
    wet = counts[solvent_kind]
    if wet <= 0:
        raise ValueError("solvent evaporation: the composition contains no solvent")
    block_start = sum(counts[:solvent_kind])
    dry = _js_round_scalar(wet * dwc["residualSolventFraction"])
    if dry >= wet:
        raise ValueError(
            f"solvent evaporation: residualSolventFraction={dwc['residualSolventFraction']} removes not a single bead "
            f"(wet {wet}, dry {dry}) -- the dry phase must remove the solvent")
    box = tuple(float(x) for x in box)
    dry_box = computeDryBox(box, block_start + dry, dwc["targetDryDensity"])
    max_radius = max(m["radiusSigma"] for m in soup["monomers"])
    solvent_radius = soup["monomers"][solvent_kind]["radiusSigma"]
    stiffest = wcaCutoff(p["sigma"] * (solvent_radius + max_radius) * 0.5)
    d0 = (soup.get("coldStartRelax") or {}).get("maxDisplacementSigma", 0.1) * p["sigma"]
    wet_volume = box[0] * box[1] * box[2]
    mean_spacing = float(np.cbrt(wet_volume / (block_start + wet)))
    need = max(0.0, stiffest - dwc["insertionMinSeparationSigma"] * p["sigma"]) + mean_spacing
    return EvaporationPlan(
        solventKind=solvent_kind, solventBlockStart=block_start, wetSolventCount=wet, drySolventCount=dry,
        organicCount=block_start, wetBox=box, dryBox=dry_box, rampSteps=int(dwc["evaporationRampSteps"]),
        minSeparationSigma=float(dwc["insertionMinSeparationSigma"]),
        relaxIterations=max(1, math.ceil((2 * need) / d0)),
        concentrationFactor=wet_volume / (dry_box[0] * dry_box[1] * dry_box[2]))


def evaporationLadder(plan: EvaporationPlan, targetDryDensity: float) -> List[Dict[str, Any]]:
    """soup-evaporate-plan.ts:132-171: rampSteps (box, solvent) rungs, the box uniform in ln L and the
    solvent count following a log-linear ramp of the total density, clamped monotone inside the
    pool; the last rung is exactly the dry box and the residual count."""
    wet_box, dry_box = plan.wetBox, plan.dryBox
    vol = lambda b: b[0] * b[1] * b[2]  # noqa: E731
    n_wet = plan.organicCount + plan.wetSolventCount
    rho_wet = n_wet / vol(wet_box)
    ln_box = math.log(dry_box[0] / wet_box[0])
    ln_rho = math.log(targetDryDensity / rho_wet)
    out: List[Dict[str, Any]] = []
    for s in range(1, plan.rampSteps + 1):
        if s == plan.rampSteps:
            out.append({"box": dry_box, "solvent": plan.drySolventCount})
            continue
        f = s / plan.rampSteps
        scale = math.exp(ln_box * f)
        b = (wet_box[0] * scale, wet_box[1] * scale, wet_box[2] * scale)
        rho = rho_wet * math.exp(ln_rho * f)
        solvent = _js_round_scalar(rho * vol(b)) - plan.organicCount
        prev = out[-1]["solvent"] if out else plan.wetSolventCount
        out.append({"box": b, "solvent": min(prev, max(plan.drySolventCount, solvent))})
    return out


def rehydrationLadder(plan: EvaporationPlan) -> List[Dict[str, Any]]:
    """soup-evaporate-plan.ts:183-198: the box ladder walked back with the solvent held at the
    residual count, the whole pool returning on the last rung at the wet box."""
    wet_box, dry_box = plan.wetBox, plan.dryBox
    ln_box = math.log(wet_box[0] / dry_box[0])
    out: List[Dict[str, Any]] = []
    for s in range(1, plan.rampSteps + 1):
        if s == plan.rampSteps:
            out.append({"box": wet_box, "solvent": plan.wetSolventCount})
            continue
        scale = math.exp((ln_box * s) / plan.rampSteps)
        out.append({"box": (dry_box[0] * scale, dry_box[1] * scale, dry_box[2] * scale), "solvent": plan.drySolventCount})
    return out


def sampleInsertionPositions(active: np.ndarray, activeN: int, box: Sequence[float], count: int, minSep: float,
                             rng_state: int) -> Dict[str, Any]:
    """sampleInsertionPositions (soup-evaporate-plan.ts:202-304) through the library's host code.
    ``rng_state`` is the mulberry32 state the candidates are drawn from; the result carries the
    advanced state as ``rng``."""
    lib = native.load_library()
    act = np.ascontiguousarray(active, dtype=np.float32).reshape(-1)
    if act.size < 4 * activeN:
        raise ValueError("sampleInsertionPositions: active holds fewer than activeN particles")
    bx = (C.c_double * 3)(*[float(x) for x in box])
    st = C.c_uint32(int(rng_state) & 0xFFFFFFFF)
    out = np.zeros(4 * int(count), dtype=np.float32)
    short = C.c_uint32()
    mn = C.c_double()
    r = lib.soup_sample_insertion(act.ctypes.data_as(C.c_void_p), C.c_uint32(int(activeN)), bx, C.c_uint32(int(count)),
                                  C.c_double(float(minSep)), C.byref(st), out.ctypes.data_as(C.c_void_p), C.byref(short),
                                  C.byref(mn))
    if r != 0:
        raise ValueError("sampleInsertionPositions: invalid box, separation or non-finite active positions")
    return {"positions": out, "shortOfFloor": int(short.value), "minAchieved": float(mn.value), "rng": int(st.value)}


# ---- the live half ---------------------------------------------------------------------------------
def _state(sys_) -> Dict[str, np.ndarray]:
    """Every per-particle array soup_reset_particles takes, in particle order."""
    rng = sys_.rngState()
    return {"positions": sys_.particles(), "velocities": sys_.velocities(), "bondSlots": sys_.bondSlots(),
            "centerLink": sys_.centerLinks(), "centerHeldSteps": sys_.centerHeldSteps(), "frozen": sys_.frozen(),
            "bondRng": rng["bond"], "thermoRng": rng["thermo"], "charges": sys_.charges()}


_PER = {"positions": 4, "velocities": 4, "bondSlots": 3, "centerLink": 1, "centerHeldSteps": 1, "frozen": 1,
        "bondRng": 1, "thermoRng": 1, "charges": 1}


def _reset(sys_, n: int, st: Dict[str, np.ndarray]) -> None:
    """soup_reset_particles with the first n particles of st, then the host's live count."""
    keep = {}
    for k, per in _PER.items():
        a = np.ascontiguousarray(st[k])
        if a.size < per * n:
            raise ValueError(f"reset: {k} holds {a.size // per} particles, {n} requested")
        keep[k] = np.ascontiguousarray(a[: per * n])
    s = native.SoupInitState()
    s.positions = keep["positions"].ctypes.data
    s.velocities = keep["velocities"].ctypes.data
    s.bond_slots = keep["bondSlots"].ctypes.data
    s.center_link = keep["centerLink"].ctypes.data
    s.center_held = keep["centerHeldSteps"].ctypes.data
    s.frozen = keep["frozen"].ctypes.data
    s.bond_rng = keep["bondRng"].ctypes.data
    s.thermo_rng = keep["thermoRng"].ctypes.data
    s.charges = keep["charges"].ctypes.data
    sys_._call(sys_._lib.soup_reset_particles, C.c_uint32(int(n)), C.byref(s))
    R = sys_._R
    R.N = int(n)
    R.cfg.n = int(n)
    sys_.N = int(n)


def setActiveCountDEBUG(sys_, n: int) -> None:
    """The live count set to n <= the current one by truncation (the reset path at a fixed box)."""
    _reset(sys_, int(n), _state(sys_))


def evaporateSolventTo(sys_, plan: EvaporationPlan, targetSolvent: int) -> None:
    """soup-evaporate.ts:84-89: the live count shrinks to solventBlockStart + targetSolvent at the
    current box (removal creates no overlap, so nothing is relaxed)."""
    live = sys_.N - plan.solventBlockStart
    if targetSolvent > live:
        raise ValueError(f"evaporation: {targetSolvent} solvent beads requested, but {live} are live now")
    _reset(sys_, plan.solventBlockStart + int(targetSolvent), _state(sys_))


def _max_abs(f: np.ndarray, n: int) -> float:
    """maxAbs / maxAbsForce (soup-evaporate.ts:325-335): infinity on any non-finite component."""
    f3 = np.asarray(f, dtype=np.float32).reshape(-1, 4)[:n, :3]
    if not np.isfinite(f3).all():
        return math.inf
    return float(np.abs(f3).max(initial=0.0))


def _d0(sys_) -> float:
    R = sys_._R
    return (R.soup.get("coldStartRelax") or {}).get("maxDisplacementSigma", 0.1) * R.p["sigma"]


def _relax(sys_, iterations: int, d0: float) -> None:
    """relaxIterations (soup/src/soup-relax.ts:87-135): normalised descent with a linearly decaying
    cap, rebuilding the lists every iteration, then a final rebuild and force."""
    sys_._call(sys_._lib.soup_relax, C.c_uint32(int(iterations)), C.c_double(float(d0)))


def rehydrateSolventTo(sys_, plan: EvaporationPlan, targetSolvent: int, newBox: Sequence[float], seed: int) -> Dict[str, Any]:
    """soup-evaporate.ts:140-244: targetSolvent - live solvent beads return at positions from the
    insertion sampler, with fresh Maxwell-Boltzmann velocities, per-index RNG seeds mixed with the
    step counter, empty bond slots and links; then the minimiser runs relaxIterations with every
    particle mobile, and the displacement it caused on the particles already present is reported."""
    R = sys_._R
    frozen = int((sys_.frozen() != 0).sum())
    if frozen > 0:
        raise ValueError(f"rehydration: the system carries {frozen} immobile beads -- a cycle with evaporation requires "
                         "clay=false/frozenBulkCatalysts=0")
    live = sys_.N - plan.solventBlockStart
    count = int(targetSolvent) - live
    if count < 0:
        raise ValueError(f"rehydration: the target {targetSolvent} is less than the live {live}")
    insert_start = sys_.N
    box = tuple(float(x) for x in newBox)
    st = _state(sys_)
    if count == 0:
        _reset(sys_, sys_.N, st)
        return {"inserted": 0, "shortOfFloor": 0, "minAchieved": 0.0, "relaxIterations": 0, "maxForceBefore": 0.0,
                "maxForceAfter": 0.0, "preexistingRmsDisplacement": 0.0, "preexistingMaxDisplacement": 0.0,
                "displacementBound": 0.0}
    active = st["positions"]
    gs = sys_.steps
    a = ((int(seed) & 0xFFFFFFFF) ^ (((gs + 1) * 0x9E3779B9) & 0xFFFFFFFF)) & 0xFFFFFFFF
    sampled = sampleInsertionPositions(active, insert_start, box, count, plan.minSeparationSigma * R.p["sigma"], a)
    pos = sampled["positions"].reshape(count, 4).copy()
    pos[:, 3] = plan.solventKind
    # the velocity draws continue the sampler's stream: per bead x, y, z, each a (u1, u2) pair
    u = Mulberry32(sampled["rng"]).take(6 * count).astype(np.float64).reshape(count, 3, 2)
    s = math.sqrt(R.p["thermostat"]["kT"])
    vel = np.zeros((count, 4), dtype=np.float32)
    vel[:, :3] = (s * _bm(u[:, :, 0], u[:, :, 1])).astype(np.float32)
    i1 = np.arange(insert_start + 1, insert_start + count + 1, dtype=np.uint64)
    sd, g32 = np.uint64(int(seed) & 0xFFFFFFFF), np.uint64(gs & 0xFFFFFFFF)
    m32 = np.uint64(0xFFFFFFFF)
    bond_rng = (sd ^ ((i1 * np.uint64(2654435761)) & m32) ^ np.uint64(0x9E3779B9) ^ g32).astype(np.uint32)
    thermo_rng = (sd ^ ((i1 * np.uint64(3266489917)) & m32) ^ np.uint64(0x85EBCA6B) ^ g32).astype(np.uint32)
    new = {
        "positions": np.concatenate([active, pos.reshape(-1)]),
        "velocities": np.concatenate([st["velocities"], vel.reshape(-1)]),
        "bondSlots": np.concatenate([st["bondSlots"], np.full(3 * count, NONE_U32, np.uint32)]),
        "centerLink": np.concatenate([st["centerLink"], np.full(count, NONE_U32, np.uint32)]),
        "centerHeldSteps": np.concatenate([st["centerHeldSteps"], np.zeros(count, np.uint32)]),
        "frozen": np.concatenate([st["frozen"], np.zeros(count, np.uint32)]),
        "bondRng": np.concatenate([st["bondRng"], bond_rng]),
        "thermoRng": np.concatenate([st["thermoRng"], thermo_rng]),
        "charges": np.concatenate([st["charges"], np.zeros(count, np.float32)]),
    }
    _reset(sys_, plan.solventBlockStart + int(targetSolvent), new)
    before = _max_abs(sys_.forces(), sys_.N)
    d0 = _d0(sys_)
    _relax(sys_, plan.relaxIterations, d0)
    after_f = _max_abs(sys_.forces(), sys_.N)
    after = sys_.particles().reshape(-1, 4)[:insert_start, :3].astype(np.float64)
    prev = active.reshape(-1, 4)[:insert_start, :3].astype(np.float64)
    b = np.asarray(box, np.float64)
    d = after - prev
    r = np.floor(d / b)
    m = d - (r + ((d / b) - r >= 0.5)) * b
    d2 = (m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1]) + m[:, 2] * m[:, 2]
    sum_sq = float(np.cumsum(d2)[-1])  # summed in index order, as the reference's loop does
    return {"inserted": count, "shortOfFloor": sampled["shortOfFloor"], "minAchieved": sampled["minAchieved"],
            "relaxIterations": plan.relaxIterations, "maxForceBefore": before, "maxForceAfter": after_f,
            "preexistingRmsDisplacement": math.sqrt(sum_sq / max(1, insert_start)),
            "preexistingMaxDisplacement": math.sqrt(float(d2.max(initial=0.0))),
            "displacementBound": (d0 * (plan.relaxIterations + 1)) / 2}


def desorbOverstretchedTethers(sys_, box: Sequence[float], maxLength: float) -> Dict[str, Any]:
    """soup-evaporate.ts:280-314: every centre link at least maxLength long (minimum image in `box`)
    is cleared on both ends and counted into the stretch desorption counter; when any is cleared the
    hold clocks of all live particles are reset, as the reference writes them."""
    from .box_scale import mi3Distance

    links = sys_.centerLinks()
    pos = sys_.particles()
    n = sys_.N
    i = np.nonzero((links != NONE_U32) & (links < n))[0]
    owner = links[i].astype(np.int64)
    r = mi3Distance(pos, box, i, owner) if i.size else np.zeros(0)
    longest = float(r.max(initial=0.0))
    hit = r >= maxLength
    cleared = int(hit.sum())
    if cleared > 0:
        nxt = links.copy()
        nxt[i[hit]] = NONE_U32
        nxt[owner[hit]] = NONE_U32
        sys_._write("CENTER_LINK", nxt)
        sys_._write("CENTER_HELD", np.zeros(n, np.uint32))
        d = sys_._read("DESORB", np.uint32).copy()
        d[0] = (int(d[0]) + cleared) & 0xFFFFFFFF
        sys_._write("DESORB", d)
    return {"cleared": cleared, "maxLength": float(maxLength), "longest": longest}


def applyEvaporatingTransition(sys_, plan: EvaporationPlan, phase: str, log=None) -> List[Dict[str, Any]]:
    """soup-box-scale.ts:375-450: one wet <-> dry transition as a ladder of (live count, box) rungs.
    Drying removes solvent and then contracts; rehydrating expands and then inserts. After each rung
    the over-stretched tethers desorb, the force is recomputed (F(x) of the new geometry) and, when
    its largest component exceeds sigma / (13 dt^2), the minimiser runs relaxIterations; each rung
    but the last is followed by rampRelaxSteps ordinary steps."""
    from .box_scale import applyBoxScaleOnce

    R = sys_._R
    dwc = R.soup["dryWetCycle"]
    ladder = evaporationLadder(plan, dwc["targetDryDensity"]) if phase == "dry" else rehydrationLadder(plan)
    reports: List[Dict[str, Any]] = []
    d0 = _d0(sys_)
    sigma, dt = R.p["sigma"], R.p["integrator"]["dt"]
    lam = math.exp(abs(math.log(plan.wetBox[0] / plan.dryBox[0])) / len(ladder))
    for s, rung in enumerate(ladder):
        if phase == "dry":
            evaporateSolventTo(sys_, plan, rung["solvent"])
            applyBoxScaleOnce(sys_, rung["box"])
        else:
            applyBoxScaleOnce(sys_, rung["box"])
            if rung["solvent"] > sys_.N - plan.solventBlockStart:
                reports.append(rehydrateSolventTo(sys_, plan, rung["solvent"], rung["box"], R.seed))
        tether = desorbOverstretchedTethers(sys_, rung["box"], R.p["fene"]["rInf"] / lam)
        spike = _max_abs(sys_.forces(), sys_.N)
        relaxed = 0
        if spike > sigma / (13 * dt * dt):
            relaxed = plan.relaxIterations
            _relax(sys_, relaxed, d0)
        after = _max_abs(sys_.forces(), sys_.N) if relaxed > 0 else spike
        if log is not None:
            log({"increment": s + 1, "of": len(ladder), "box": tuple(rung["box"]), "solvent": sys_.N - plan.solventBlockStart,
                 "N": sys_.N, "maxForce": spike, "relaxIterations": relaxed, "maxForceAfter": after, "tether": tether})
        if s < len(ladder) - 1 and dwc["rampRelaxSteps"] > 0:
            sys_.step(dwc["rampRelaxSteps"])
    return reports

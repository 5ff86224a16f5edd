"""Grid / Verlet-list planning (pure, device-free).

Restates soup/src/soup-plan.ts:88-324: ``planSoupGrid`` (:112-149), ``deriveGridGeometry``
(:170-235, with its three throwing completeness guards), ``deriveListCapacity`` (:258-280),
``densestDensityOf`` (:295-318) and ``verletListBytes`` (:322-324). Error messages keep the
reference's wording so callers matching on them keep working.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

from .rules import loadParams, loadSoup, wcaCutoff


@dataclass
class SoupPlan:
    N: int
    box: Tuple[float, float, float]
    cellSize: float
    effectiveWalkRadius: int
    dims: Tuple[int, int, int]
    ncells: int
    minCells: int
    valid: bool
    reason: Optional[str]


def _js_fixed(x: float, digits: int) -> str:
    return f"{x:.{digits}f}"


def _fmt_num(x: float) -> str:
    """JS Number -> string for integral/simple values used in messages."""
    if float(x).is_integer():
        return str(int(x))
    return repr(float(x))


def planSoupGrid(box, startCounts: Optional[Dict[str, int]] = None) -> SoupPlan:
    soup = loadSoup()
    p = loadParams()
    counts = dict(soup["start"])
    counts.update(startCounts or {})
    N = sum(int(counts.get(m["id"], 0)) for m in soup["monomers"])
    max_radius = max(m["radiusSigma"] for m in soup["monomers"])
    interaction_range = wcaCutoff(p["sigma"] * max_radius) + p["attraction"]["wc"]
    cell_size = interaction_range / soup["neighborGrid"]["cellDivisor"]
    walk_radius = math.ceil(interaction_range / cell_size)
    verlet = soup["verletList"]
    list_range = interaction_range + verlet["skin"]
    list_walk = math.ceil(list_range / cell_size)
    eff = list_walk if verlet["enabled"] else walk_radius
    dims = tuple(max(1, int(math.floor(box[a] / cell_size))) for a in range(3))
    ncells = dims[0] * dims[1] * dims[2]
    min_cells = 2 * eff + 1
    grid_ok = all(d >= min_cells for d in dims)
    mi_ok = min(box) / 2 > p["bend"]["r0"]
    valid = grid_ok and mi_ok
    reason = None
    if not valid:
        reason = (
            f"neighbour grid: box=[{_fmt_num(box[0])},{_fmt_num(box[1])},{_fmt_num(box[2])}] gives cellSize={_js_fixed(cell_size, 4)}, "
            f"dims=[{dims[0]},{dims[1]},{dims[2]}] and min(box)/2={_js_fixed(min(box) / 2, 4)}; it needs dims>={min_cells} "
            f"(2*effectiveWalkRadius+1) on all three axes and min(box)/2 > bend.r0={_fmt_num(p['bend']['r0'])}"
        )
    return SoupPlan(N, tuple(box), cell_size, eff, dims, ncells, min_cells, valid, reason)


@dataclass
class GridGeometry:
    interactionRange: float
    cellSize: float
    walkRadius: int
    listRange: float
    listBuildWalkRadius: int
    effectiveWalkRadius: int


def deriveGridGeometry(soup: dict, p: dict, kT: float) -> GridGeometry:
    max_radius = max(m["radiusSigma"] for m in soup["monomers"])
    interaction_range = wcaCutoff(p["sigma"] * max_radius) + p["attraction"]["wc"]
    cell_size = interaction_range / soup["neighborGrid"]["cellDivisor"]
    walk_radius = math.ceil(interaction_range / cell_size)
    if walk_radius * cell_size < interaction_range - 1e-6:
        raise ValueError(
            f"neighbour grid: walkRadius={walk_radius} * cellSize={cell_size:.6f} = {walk_radius * cell_size:.6f} "
            f"does not cover interactionRange={interaction_range:.6f}, so the traversal completeness guarantee is broken"
        )
    verlet = soup["verletList"]
    list_range = interaction_range + verlet["skin"]
    list_walk = math.ceil(list_range / cell_size)
    if list_walk * cell_size < list_range - 1e-6:
        raise ValueError(
            f"Verlet list: listBuildWalkRadius={list_walk} * cellSize={cell_size:.6f} = {list_walk * cell_size:.6f} "
            f"does not cover listRange={list_range:.6f} (interactionRange+skin), so the traversal completeness guarantee is broken"
        )
    if verlet["enabled"]:
        v_bound = 2 * math.sqrt(3 * kT)
        drift_bound = 2 * verlet["rebuildEvery"] * p["integrator"]["dt"] * v_bound
        if drift_bound > verlet["skin"]:
            raise ValueError(
                f"Verlet list: 2*rebuildEvery*dt*vBound={drift_bound:.4f} exceeds skin={_fmt_num(verlet['skin'])} "
                f"at kT={_fmt_num(kT)}: the rebuild is not frequent enough (or skin is too small) for this temperature"
            )
    eff = list_walk if verlet["enabled"] else walk_radius
    return GridGeometry(interaction_range, cell_size, walk_radius, list_range, list_walk, eff)


def deriveListCapacity(soup: dict, listRange: float, capacityN: int, densestDensity: float) -> dict:
    vl = soup["verletList"]
    f = vl["capacitySafetyFactor"]
    if not (capacityN > 0) or not (densestDensity > 0) or not (f > 0):
        return {"capacity": vl["listCapacity"], "uniform": 0.0, "safetyFactor": f, "derived": False}
    uniform = (4 * math.pi / 3) * listRange ** 3 * densestDensity
    capacity = max(64, min(capacityN, max(vl["capacityFloor"], math.ceil(f * uniform))))
    return {"capacity": capacity, "uniform": uniform, "safetyFactor": f, "derived": True}


def densestDensityOf(soup: dict, capacityN: int, creationBox, liveN: int, liveBox, cycling: bool) -> float:
    vol = lambda b: b[0] * b[1] * b[2]  # noqa: E731
    return max(capacityN / vol(creationBox), liveN / vol(liveBox), soup["dryWetCycle"]["targetDryDensity"] if cycling else 0.0)


def verletListBytes(capacityN: int, listCapacity: int) -> int:
    return capacityN * listCapacity * 4

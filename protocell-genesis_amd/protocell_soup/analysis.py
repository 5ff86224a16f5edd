"""Snapshot analysis of a soup: what soup/src/stages.ts:204-289 (detectStage) measures every sample.

The graph and grid work runs on the GPU through the analysis context of the C-ABI:
- bond-graph components for findAmphiphiles;
- member connectivity for clusterComponents;
- the periodic encapsulated-water flood;
- the open-boundary cavity flood.
The per-aggregate geometry (unwrap, gyration tensor, head-density profiles) is a few FP64
reductions over one aggregate's members and runs here, in the order and precision the reference
uses. Float32Array stores in the reference are float32 stores here; JavaScript number arithmetic is
float64; Math.round is floor(x + 0.5) semantics.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import native
from .rules import _load_json, loadParams, loadSoup, wcaCutoff

NONE_U32 = 0xFFFFFFFF
HEAD_PEAKS_UNAVAILABLE = "unavailable"


def _js_round(x):
    """Math.round for float64 arrays or scalars: round half toward +infinity."""
    f = np.floor(x)
    return f + ((x - f) >= 0.5)


def _mi(d, box):
    return d - _js_round(d / box) * box


def _wrap(v, box):
    return v - np.floor(v / box) * box


def _seqsum(a: np.ndarray) -> float:
    """Left-to-right float64 sum (JavaScript's `s += x` loop), not numpy's pairwise sum."""
    return float(np.cumsum(a, dtype=np.float64)[-1]) if a.size else 0.0


# ---------------------------------------------------------------------------------------------
# The device context
# ---------------------------------------------------------------------------------------------
class AnalysisContext:
    """One GPU snapshot (analysis context of the C-ABI). Not safe for concurrent use."""

    def __init__(self):
        self._lib = native.load_library()
        self._a = C.c_void_p()
        code = self._lib.soup_analysis_create(C.byref(self._a))
        if code != 0:
            msg = self._lib.soup_analysis_last_error(None)
            raise native.SoupError(msg.decode() if msg else f"soup_analysis_create failed ({code})", code)
        self.n = 0

    def _check(self, code: int) -> None:
        if code != 0:
            msg = self._lib.soup_analysis_last_error(self._a)
            raise native.SoupError(msg.decode() if msg else f"analysis call failed ({code})", code)

    def close(self) -> None:
        if self._a:
            self._lib.soup_analysis_destroy(self._a)
            self._a = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, particles: np.ndarray, bond_slots: np.ndarray) -> None:
        pos = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
        n = pos.size // 4
        bs = np.ascontiguousarray(bond_slots, dtype=np.uint32).reshape(-1)
        if bs.size != 3 * n:
            raise ValueError(f"analysis: {bs.size} bond-slot entries for {n} particles (need 3 per particle)")
        self._check(self._lib.soup_analysis_load(self._a, C.c_uint32(n), pos.ctypes.data, bs.ctypes.data))
        self.n = n

    def load_system(self, sys) -> None:
        self._check(self._lib.soup_analysis_load_system(self._a, sys._h))
        self.n = sys.N

    def amphiphiles(self, polar_mask: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        na, nc = C.c_uint32(), C.c_uint32()
        self._check(self._lib.soup_analysis_amphiphiles(self._a, C.c_uint32(polar_mask), C.byref(na), C.byref(nc)))
        heads = np.zeros(max(1, na.value), np.uint32)
        off = np.zeros(na.value + 1, np.uint32)
        chains = np.zeros(max(1, nc.value), np.uint32)
        tails = np.zeros(max(2, 2 * na.value), np.uint32)
        self._check(self._lib.soup_analysis_amphiphiles_read(self._a, heads.ctypes.data, off.ctypes.data, chains.ctypes.data,
                                                             tails.ctypes.data))
        return heads[: na.value], off, chains[: nc.value], tails[: 2 * na.value].reshape(-1, 2)

    def clusters(self, members: np.ndarray, box, cutoff: float, periodic_z: bool) -> np.ndarray:
        mem = np.ascontiguousarray(members, dtype=np.uint32)
        labels = np.zeros(max(1, mem.size), np.uint32)
        bx = (C.c_double * 3)(*box)
        self._check(self._lib.soup_analysis_clusters(self._a, mem.ctypes.data, C.c_uint32(mem.size), bx, C.c_double(cutoff),
                                                     C.c_int(1 if periodic_z else 0), labels.ctypes.data))
        return labels[: mem.size]

    def water(self, members: np.ndarray, water_kind: int, box, centre, cell: float, radius: float) -> native.SoupWaterCounts:
        mem = np.ascontiguousarray(members, dtype=np.uint32)
        out = native.SoupWaterCounts()
        self._check(self._lib.soup_analysis_water(self._a, mem.ctypes.data, C.c_uint32(mem.size), C.c_uint32(water_kind),
                                                  (C.c_double * 3)(*box), (C.c_double * 3)(*centre), C.c_double(cell),
                                                  C.c_double(radius), C.byref(out)))
        return out

    def cavity(self, positions: np.ndarray, box, cell: float, radius: float) -> int:
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        out = C.c_uint64()
        self._check(self._lib.soup_analysis_cavity(self._a, pos.ctypes.data, C.c_uint32(pos.size // 4), (C.c_double * 3)(*box),
                                                   C.c_double(cell), C.c_double(radius), C.byref(out)))
        return int(out.value)


# ---------------------------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------------------------
def polarMask(monomers: Sequence[dict]) -> int:
    mask = 0
    for k, m in enumerate(monomers):
        if m.get("polar"):
            mask |= 1 << k
    return mask


def derivedMinAmphiphilesPerAggregate() -> int:
    """soup/src/aggregates.ts:180-196."""
    p = loadParams()
    soup = loadSoup()
    carbon = next((m for m in soup["monomers"] if m["kind"] == "carbon"), None)
    if carbon is None:
        raise ValueError("data/soup.json: no monomer of kind carbon found")
    tail = wcaCutoff(p["sigma"] * carbon["radiusSigma"])
    gate = next((g for g in json.loads(_load_json("literature.json"))["gates"] if g["id"] == "area-per-lipid"), None)
    if gate is None or "min" not in gate["target"] or "max" not in gate["target"]:
        raise ValueError("data/literature.json: missing area-per-lipid.target.min/max")
    mid = (gate["target"]["min"] + gate["target"]["max"]) / 2
    return math.ceil(4 * math.pi * tail * tail / mid)


def loadStageThresholds() -> Dict[str, Any]:
    """soup/src/stages.ts:109-140, with both consistency checks."""
    t = loadSoup().get("stageThresholds")
    if not t:
        raise ValueError("data/soup.json: missing field stageThresholds")
    gate = next((g for g in json.loads(_load_json("literature.json"))["gates"] if g["id"] == "closure"), None)
    if gate is None or "min" not in gate["target"]:
        raise ValueError("data/literature.json: missing closure.target.min")
    if t["enclosedVolume"] != gate["target"]["min"]:
        raise ValueError(f"data/soup.json stageThresholds.enclosedVolume ({t['enclosedVolume']}) does not match "
                         f"data/literature.json closure.target.min ({gate['target']['min']})")
    derived = derivedMinAmphiphilesPerAggregate()
    if t["minAmphiphilesPerAggregate"] != derived:
        raise ValueError(f"data/soup.json stageThresholds.minAmphiphilesPerAggregate ({t['minAmphiphilesPerAggregate']}) "
                         f"does not match the derived formula ({derived})")
    return t


def circularAlpha() -> float:
    return float(json.loads(_load_json("periodic-measurement.json"))["circularConcentration"]["alpha"])


def memberCutoff(soup=None, p=None) -> float:
    """detectStage's aggregate cutoff (stages.ts:226-229): wcaCutoff(sigma * max member radius) + wc."""
    soup = soup or loadSoup()
    p = p or loadParams()
    radii = [m["radiusSigma"] for m in soup["monomers"] if m["kind"] in ("carbon", "head")]
    return wcaCutoff(p["sigma"] * max(radii)) + p["attraction"]["wc"]


# ---------------------------------------------------------------------------------------------
# findAmphiphiles (soup/src/amphiphile.ts:55-141)
# ---------------------------------------------------------------------------------------------
def bondSlotsFromPairs(n: int, bonds: np.ndarray) -> np.ndarray:
    """Three slots per particle from flat (i, j) pairs, filled in pair order. For pairs in bonds()
    order (ascending lower index, then slot) the device adjacency equals amphiphile.ts:66-73's."""
    slots = np.full((n, 3), NONE_U32, np.uint32)
    fill = np.zeros(n, np.int64)
    b = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
    for i, j in b:
        for x, y in ((i, j), (j, i)):
            if fill[x] < 3:
                slots[x, fill[x]] = y
            fill[x] += 1
    if (fill > 3).any():
        # a bead with more than three bonds can never be part of an amphiphile (degree <= 2);
        # three recorded slots keep it rejected
        pass
    return slots.reshape(-1)


def _amphiphile_list(heads, off, chains, tails) -> List[Dict[str, Any]]:
    out = []
    for a in range(heads.size):
        chain = [int(x) for x in chains[off[a]:off[a + 1]]]
        tl = [int(tails[a, 0])] + ([int(tails[a, 1])] if tails[a, 1] else [])
        out.append({"headIndex": int(heads[a]), "chain": chain, "length": len(chain), "tailLengths": tl})
    return out


def findAmphiphiles(particles: np.ndarray, bonds: np.ndarray, monomers: Sequence[dict],
                    ctx: Optional[AnalysisContext] = None) -> List[Dict[str, Any]]:
    """findAmphiphiles(particles, bonds, monomers) on host arrays (uploaded to the GPU context)."""
    ctx = ctx or AnalysisContext()
    pos = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
    ctx.load(pos, bondSlotsFromPairs(pos.size // 4, bonds))
    return _amphiphile_list(*ctx.amphiphiles(polarMask(monomers)))


def findAmphiphilesLoaded(ctx: AnalysisContext, monomers: Sequence[dict]) -> List[Dict[str, Any]]:
    """findAmphiphiles over the snapshot already loaded in ``ctx``."""
    return _amphiphile_list(*ctx.amphiphiles(polarMask(monomers)))


def amphiphileHistogram(amphiphiles: Sequence[dict]) -> Dict[int, int]:
    hist: Dict[int, int] = {}
    for a in amphiphiles:
        hist[a["length"]] = hist.get(a["length"], 0) + 1
    return hist


def memberIndicesOf(amphiphiles: Sequence[dict]) -> List[int]:
    """soup/src/aggregates.ts:122-129: heads and chains, in Set insertion order."""
    seen = {}
    for a in amphiphiles:
        seen.setdefault(a["headIndex"], None)
        for c in a["chain"]:
            seen.setdefault(c, None)
    return list(seen)


# ---------------------------------------------------------------------------------------------
# Components (engine/src/aggregate.ts:65-212)
# ---------------------------------------------------------------------------------------------
def componentsFromLabels(labels: np.ndarray) -> List[np.ndarray]:
    """Local-index lists, in first-occurrence order, then stably sorted by size, largest first
    (aggregate.ts:201-212). ``labels`` are the smallest local index of each component."""
    if labels.size == 0:
        return []
    roots, first = np.unique(labels, return_index=True)  # first occurrence = the root itself
    order = np.argsort(labels, kind="stable")
    counts = np.bincount(np.searchsorted(roots, labels), minlength=roots.size)
    starts = np.concatenate(([0], np.cumsum(counts)))
    groups = [order[starts[k]:starts[k + 1]] for k in range(roots.size)]  # ascending local indices
    rank = sorted(range(roots.size), key=lambda k: (-int(counts[k]), int(first[k])))
    return [groups[k] for k in rank]


def clusterComponents(ctx: AnalysisContext, members: Sequence[int], box, cutoff: float, periodicZ: bool = True):
    return componentsFromLabels(ctx.clusters(np.asarray(members, dtype=np.uint32), box, cutoff, periodicZ))


# ---------------------------------------------------------------------------------------------
# Periodic centre and the encapsulated-water flood (soup/src/water-closure.ts)
# ---------------------------------------------------------------------------------------------
def periodicCentreOf(particles: np.ndarray, idx: Sequence[int], box, alpha: float) -> Optional[List[float]]:
    """water-closure.ts:56-100: circular mean per axis; None when the Rayleigh test refuses an axis."""
    pos = np.asarray(particles, dtype=np.float32).reshape(-1, 4)
    sel = np.asarray(idx, dtype=np.int64)
    n = sel.size
    if n == 0:
        raise ValueError("circularAxisStat: empty list of coordinates")
    out = []
    for a in range(3):
        L = float(box[a])
        theta = (2 * math.pi * _wrap(pos[sel, a].astype(np.float64), L)) / L
        sc = _seqsum(np.cos(theta)) / n
        ss = _seqsum(np.sin(theta)) / n
        R = math.sqrt(sc * sc + ss * ss)
        if not R >= math.sqrt(-math.log(alpha) / n):
            return None
        ang = math.atan2(ss, sc)
        out.append(float(_wrap((ang / (2 * math.pi)) * L, L)))
    return out


def encapsulatedWaterVolume(ctx: AnalysisContext, members: Sequence[int], water_kind: int, box, centre, cell: float,
                            radius: float, enclosedVolumeMin: float) -> Dict[str, Any]:
    """encapsulatedWaterVolume (water-closure.ts:268-363) over the loaded snapshot, the water beads
    being every bead of ``water_kind``."""
    w = ctx.water(np.asarray(members, dtype=np.uint32), water_kind, box, centre, cell, radius)
    bulk = w.reached_empty_water / (w.reached_empty_cells * cell * cell * cell) if w.reached_empty_cells > 0 else 0.0
    threshold = enclosedVolumeMin * bulk if bulk > 0 else float("nan")
    nx, ny = w.dims[0], w.dims[1]
    z = w.seed_cell // (nx * ny)
    rem = w.seed_cell - z * nx * ny
    y = rem // nx
    x = rem - y * nx
    dx = float(_mi((x + 0.5) * cell - centre[0], box[0]))
    dy = float(_mi((y + 0.5) * cell - centre[1], box[1]))
    dz = float(_mi((z + 0.5) * cell - centre[2], box[2]))
    return {
        "encapsulatedCount": int(w.encapsulated),
        "totalWater": int(w.total_water),
        "bulkWaterDensity": bulk,
        "encapsulationThresholdCount": threshold,
        "closed": bool(math.isfinite(threshold) and w.encapsulated > threshold),
        "seedDistFromCentre": math.sqrt(dx * dx + dy * dy + dz * dz),
        "theoreticalMaxDist": (math.sqrt(3) * min(box)) / 2,
        "totalEmptyCells": int(w.total_empty_cells),
        "unreachedCells": int(w.unreached_cells),
    }


# ---------------------------------------------------------------------------------------------
# Per-aggregate geometry (engine/src/aggregate.ts:225-362, engine/src/metrics.ts, aggregates.ts)
# ---------------------------------------------------------------------------------------------
def unwrapAggregate(raw: np.ndarray, box) -> np.ndarray:
    """aggregate.ts:237-253 with all three axes periodic: f32(ref + mi(v - ref))."""
    p = np.asarray(raw, dtype=np.float32).reshape(-1, 4)
    out = np.empty_like(p)
    if p.shape[0] == 0:
        return out.reshape(-1)
    for a in range(3):
        ref = float(p[0, a])
        v = p[:, a].astype(np.float64)
        out[:, a] = (ref + _mi(v - ref, float(box[a]))).astype(np.float32)
    out[:, 3] = p[:, 3]
    return out.reshape(-1)


def _eigenvalues3(Sxx, Syy, Szz, Sxy, Sxz, Syz):
    """Smith's closed form (aggregate.ts:260-279)."""
    p1 = Sxy * Sxy + Sxz * Sxz + Syz * Syz
    if p1 < 1e-30:
        return sorted([Sxx, Syy, Szz])
    q = (Sxx + Syy + Szz) / 3
    p2 = (Sxx - q) ** 2 + (Syy - q) ** 2 + (Szz - q) ** 2 + 2 * p1
    p = math.sqrt(p2 / 6)
    Bxx, Byy, Bzz = (Sxx - q) / p, (Syy - q) / p, (Szz - q) / p
    Bxy, Bxz, Byz = Sxy / p, Sxz / p, Syz / p
    det = Bxx * (Byy * Bzz - Byz * Byz) - Bxy * (Bxy * Bzz - Byz * Bxz) + Bxz * (Bxy * Byz - Byy * Bxz)
    r = max(-1.0, min(1.0, det / 2))
    phi = math.acos(r) / 3
    largest = q + 2 * p * math.cos(phi)
    smallest = q + 2 * p * math.cos(phi + (2 * math.pi) / 3)
    return sorted([smallest, 3 * q - largest - smallest, largest])


def _eigenvector3(Sxx, Syy, Szz, Sxy, Sxz, Syz, lam):
    """aggregate.ts:291-311: the largest cross product of two rows of (S - lambda I)."""
    r0, r1, r2 = (Sxx - lam, Sxy, Sxz), (Sxy, Syy - lam, Syz), (Sxz, Syz, Szz - lam)

    def cross(a, b):
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    best, best_mag = None, -1.0
    for c in (cross(r0, r1), cross(r0, r2), cross(r1, r2)):
        m = math.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
        if best is None or m > best_mag:
            best, best_mag = c, m
    if best_mag < 1e-9:
        return None
    return [best[0] / best_mag, best[1] / best_mag, best[2] / best_mag]


def shapeOf(positions: np.ndarray) -> Dict[str, Any]:
    """aggregate.ts:339-362."""
    p = np.asarray(positions, dtype=np.float32).reshape(-1, 4)
    n = p.shape[0]
    if n == 0:
        raise ValueError("shapeOf: empty set of particles")
    x, y, z = (p[:, a].astype(np.float64) for a in range(3))
    cx, cy, cz = _seqsum(x) / n, _seqsum(y) / n, _seqsum(z) / n
    dx, dy, dz = x - cx, y - cy, z - cz
    Sxx, Syy, Szz = _seqsum(dx * dx) / n, _seqsum(dy * dy) / n, _seqsum(dz * dz) / n
    Sxy, Sxz, Syz = _seqsum(dx * dy) / n, _seqsum(dx * dz) / n, _seqsum(dy * dz) / n
    moments = _eigenvalues3(Sxx, Syy, Szz, Sxy, Sxz, Syz)
    axis = _eigenvector3(Sxx, Syy, Szz, Sxy, Sxz, Syz, moments[0]) or [0.0, 0.0, 1.0]
    rg = math.sqrt(max(0.0, moments[0] + moments[1] + moments[2]))
    return {"centre": [cx, cy, cz], "radiusOfGyration": rg, "principalMoments": moments, "smallestAxis": axis}


def computeHeadPeaks(particles: np.ndarray, box, monomers: Sequence[dict], headDensityBins: int, minHeadsPerBin: float,
                     headPeakSeparationMin: float, headPeakSeparationMax: float):
    """aggregates.ts:83-120 over densityProfileZ / bilayerPeaks / subBinPeak (engine/src/metrics.ts)."""
    p = np.asarray(particles, dtype=np.float32).reshape(-1, 4)
    polar = np.array([bool(m.get("polar")) for m in monomers])
    kind = _js_round(p[:, 3].astype(np.float64)).astype(np.int64)
    is_head = np.zeros(p.shape[0], bool)
    ok = (kind >= 0) & (kind < polar.size)
    is_head[ok] = polar[kind[ok]]
    if not is_head.any():
        return 0
    z = p[:, 2].astype(np.float64)
    keep = (z >= 0) & (z < box[2])  # dropEscapedZ
    kept_heads = int((keep & is_head).sum())
    if kept_heads == 0:
        return HEAD_PEAKS_UNAVAILABLE
    if kept_heads / headDensityBins < minHeadsPerBin:
        return HEAD_PEAKS_UNAVAILABLE
    dzb = box[2] / headDensityBins
    b = np.minimum(headDensityBins - 1, np.floor(z[keep & is_head] / dzb)).astype(np.int64)
    head = np.bincount(b, minlength=headDensityBins).astype(np.float64)
    zc = (np.arange(headDensityBins) + 0.5) * dzb
    a = int(np.argmax(head))  # first maximum, as the reference's strict '>' scan
    sigma = loadParams()["sigma"]
    masked = np.where(np.abs(zc - zc[a]) <= sigma, 0.0, head)
    b2 = int(np.argmax(masked))
    if head[a] == 0 or masked[b2] == 0:
        return 1

    def sub_bin(i):
        if i == 0 or i == headDensityBins - 1 or headDensityBins < 2:
            return float(zc[i])
        denom = head[i - 1] - 2 * head[i] + head[i + 1]
        if denom >= 0:
            return float(zc[i])
        off = (0.5 * (head[i - 1] - head[i + 1])) / denom
        return float(zc[i] + max(-0.5, min(0.5, off)) * (zc[1] - zc[0]))

    za, zb = sub_bin(a), sub_bin(b2)
    sep = abs(zb - za)
    return 2 if headPeakSeparationMin <= sep <= headPeakSeparationMax else 1


def profileAlong(coords: np.ndarray, kinds: np.ndarray) -> Tuple[np.ndarray, List[float]]:
    """aggregates.ts:324-340: the coordinate (shifted to start at 0) in the z slot of a float32
    particle array, with a [1, 1, span] box."""
    c = np.asarray(coords, dtype=np.float64)
    n = c.size
    lo = float(c.min()) if n else math.inf
    hi = float(c.max()) if n else -math.inf
    span = max(1e-6, (hi - lo) * 1.0001)
    out = np.zeros((n, 4), np.float32)
    out[:, 2] = (c - lo).astype(np.float32)
    out[:, 3] = np.asarray(kinds, dtype=np.float32)
    return out.reshape(-1), [1.0, 1.0, span]


def localCavityVolume(ctx: AnalysisContext, positions: np.ndarray, cell: float, radius: float) -> float:
    """aggregates.ts:346-383: the aggregate's padded local region, flooded from its six faces."""
    p = np.asarray(positions, dtype=np.float32).reshape(-1, 4)
    if p.shape[0] == 0:
        return 0.0
    lo = [float(p[:, a].min()) for a in range(3)]
    hi = [float(p[:, a].max()) for a in range(3)]
    pad = radius * 4
    local_box = [hi[a] - lo[a] + 2 * pad for a in range(3)]
    shifted = p.copy()
    for a in range(3):
        shifted[:, a] = (p[:, a].astype(np.float64) + (pad - lo[a])).astype(np.float32)
    return ctx.cavity(shifted, local_box, cell, radius) * cell * cell * cell


def isLamellarShape(shape: Dict[str, Any], t: Dict[str, Any]) -> bool:
    return shape["flatnessRatio"] <= t["lamellarFlatnessRatio"] and shape["inPlaneSymmetry"] >= t["lamellarInPlaneSymmetryMin"]


def isBilayerShape(shape: Dict[str, Any], t: Dict[str, Any]) -> bool:
    return isLamellarShape(shape, t) and (shape["radialHeadShells"] == 2 or shape["transverseHeadShells"] == 2)


def isVesicleShape(shape: Dict[str, Any], t: Dict[str, Any]) -> bool:
    if shape["radialHeadShells"] != 2:
        return False
    if "encapsulatedWater" in shape:
        ew = shape["encapsulatedWater"]
        return ew is not None and bool(ew["closed"])
    return shape["cavityVolume"] > t["enclosedVolume"]


def _shape_of_aggregate(ctx, original_idx, amphiphile_count, particles, box, monomers, t, water_kind):
    """shapeOfAggregate (aggregates.ts:392-470)."""
    pos = np.asarray(particles, dtype=np.float32).reshape(-1, 4)
    raw = pos[np.asarray(original_idx, dtype=np.int64)]
    unwrapped = unwrapAggregate(raw, box).reshape(-1, 4)
    shape = shapeOf(unwrapped)
    l0, l1, l2 = shape["principalMoments"]
    flatness = l0 / l2 if l2 > 1e-12 else 1.0
    symmetry = l1 / l2 if l2 > 1e-12 else 1.0
    kinds = unwrapped[:, 3]
    ctr = shape["centre"]
    d = [unwrapped[:, a].astype(np.float64) - ctr[a] for a in range(3)]
    radial = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    hp = (t["aggregateHeadDensityBins"], t["aggregateMinHeadsPerBin"], t["headPeakSeparationMin"], t["headPeakSeparationMax"])
    rp, rbox = profileAlong(radial, kinds)
    radial_shells = computeHeadPeaks(rp, rbox, monomers, *hp)
    transverse_shells = 0
    if flatness <= t["lamellarFlatnessRatio"] and symmetry >= t["lamellarInPlaneSymmetryMin"]:
        ax = shape["smallestAxis"]
        trans = d[0] * ax[0] + d[1] * ax[1] + d[2] * ax[2]
        tp, tbox = profileAlong(trans, kinds)
        transverse_shells = computeHeadPeaks(tp, tbox, monomers, *hp)
    out = {
        "particleCount": int(unwrapped.shape[0]),
        "amphiphileCount": int(amphiphile_count),
        "centre": shape["centre"],
        "radiusOfGyration": shape["radiusOfGyration"],
        "principalMoments": shape["principalMoments"],
        "flatnessRatio": flatness,
        "inPlaneSymmetry": symmetry,
        "radialHeadShells": radial_shells,
        "transverseHeadShells": transverse_shells,
        "cavityVolume": localCavityVolume(ctx, unwrapped, t["closureCell"], t["closureRadius"]),
    }
    if water_kind is not None:
        centre = periodicCentreOf(particles, original_idx, box, circularAlpha())
        out["encapsulatedWater"] = (encapsulatedWaterVolume(ctx, original_idx, water_kind, box, centre, t["closureCell"],
                                                            t["closureRadius"], t["enclosedVolume"])
                                    if centre is not None else None)
    return out


def emptyAggregateAnalysis() -> Dict[str, Any]:
    return {"aggregateCount": 0, "sizeHistogram": [], "aggregates": [], "qualifyingAggregateCount": 0,
            "amphiphilesInQualifying": 0, "amphiphileShareInQualifying": 0, "hasLamellarAggregate": False,
            "hasVesicleAggregate": False}


def analyzeAggregates(ctx: AnalysisContext, particles: np.ndarray, box, monomers: Sequence[dict],
                      amphiphiles: Sequence[dict], memberIdx: Sequence[int], cutoff: float, thresholds: Dict[str, Any],
                      water_kind: Optional[int] = None) -> Dict[str, Any]:
    """analyzeAggregates (aggregates.ts:496-558) over the snapshot loaded in ``ctx`` (``particles``
    is its host copy). ``water_kind`` None = no water indices (the vacuum-cavity vesicle check)."""
    if len(memberIdx) == 0:
        return emptyAggregateAnalysis()
    idx = np.asarray(memberIdx, dtype=np.int64)
    comps = clusterComponents(ctx, idx, box, cutoff, True)
    if not comps:
        return emptyAggregateAnalysis()
    heads = np.zeros(ctx.n, bool)
    heads[[a["headIndex"] for a in amphiphiles]] = True
    prepared = []
    for local in comps:
        original = idx[local]
        prepared.append((original, int(heads[original].sum())))
    size_hist = [c for _, c in prepared]
    total = len(amphiphiles)
    qualifying = [c for _, c in prepared if c >= thresholds["minAmphiphilesPerAggregate"]]
    in_q = sum(qualifying)
    detail = min(thresholds["detailAggregateCount"], len(prepared))
    aggs = [_shape_of_aggregate(ctx, prepared[k][0], prepared[k][1], particles, box, monomers, thresholds, water_kind)
            for k in range(detail)]
    return {
        "aggregateCount": len(comps),
        "sizeHistogram": size_hist,
        "aggregates": aggs,
        "qualifyingAggregateCount": len(qualifying),
        "amphiphilesInQualifying": in_q,
        "amphiphileShareInQualifying": in_q / total if total > 0 else 0,
        "hasLamellarAggregate": any(isBilayerShape(s, thresholds) for s in aggs),
        "hasVesicleAggregate": any(isVesicleShape(s, thresholds) for s in aggs),
    }


# ---------------------------------------------------------------------------------------------
# Box-wide diagnostics and the stage ladder (soup/src/stages.ts, engine/src/closure.ts)
# ---------------------------------------------------------------------------------------------
def largestClusterCenter(ctx, particles, members, box, cutoff) -> List[float]:
    """aggregate.ts:167-195 (z open): the first largest component's minimum-image mean in x, y and
    plain mean in z."""
    idx = np.asarray(members, dtype=np.int64)
    labels = ctx.clusters(idx.astype(np.uint32), box, cutoff, False)
    roots, counts = np.unique(labels, return_counts=True)
    best = roots[int(np.argmax(counts))]  # first maximum in first-occurrence order (roots ascend)
    sel = idx[labels == best]
    p = np.asarray(particles, dtype=np.float32).reshape(-1, 4)[sel]
    x, y, z = (p[:, a].astype(np.float64) for a in range(3))
    rx, ry = x[0], y[0]
    sx = _seqsum(_mi(x - rx, box[0]))
    sy = _seqsum(_mi(y - ry, box[1]))
    sz = _seqsum(z)
    n = sel.size
    return [float(_wrap(rx + sx / n, box[0])), float(_wrap(ry + sy / n, box[1])), sz / n]


def enclosedVolumeFromPositions(ctx, particles, members, box, cell, radius) -> float:
    """closure.ts:291-323: recentre the member beads on their largest z-open cluster, then the
    open-boundary flood over the whole box."""
    p = loadParams()
    cut = wcaCutoff(p["beadSizes"]["tail_tail"]) + p["attraction"]["wc"]
    com = largestClusterCenter(ctx, particles, members, box, cut)
    shift = [box[0] / 2 - com[0], box[1] / 2 - com[1], box[2] / 2 - com[2]]
    pos = np.asarray(particles, dtype=np.float32).reshape(-1, 4)[np.asarray(members, dtype=np.int64)]
    out = pos.copy()
    out[:, 0] = _wrap(pos[:, 0].astype(np.float64) + shift[0], box[0]).astype(np.float32)
    out[:, 1] = _wrap(pos[:, 1].astype(np.float64) + shift[1], box[1]).astype(np.float32)
    out[:, 2] = (pos[:, 2].astype(np.float64) + shift[2]).astype(np.float32)
    return ctx.cavity(out, box, cell, radius) * cell * cell * cell


def stageFromEvidence(evidence: Dict[str, Any], t: Dict[str, Any]) -> str:
    """stages.ts:173-192."""
    a = evidence["aggregateAnalysis"]
    amph = evidence["amphiphileFraction"] >= t["amphiphileFraction"]
    micelle = amph and a["qualifyingAggregateCount"] >= t["minMicelleAggregates"] and \
        a["amphiphileShareInQualifying"] >= t["minAmphiphileShareInAggregates"]
    if amph and a["hasVesicleAggregate"]:
        return "vesicle"
    if amph and a["hasLamellarAggregate"]:
        return "bilayer"
    if micelle:
        return "micelles"
    if amph:
        return "amphiphiles"
    return "monomers"


def detectStage(sys, ctx: Optional[AnalysisContext] = None) -> Tuple[str, Dict[str, Any]]:
    """detectStage (stages.ts:204-289) on a live SoupSystem: the snapshot is copied to the analysis
    context on the device, the host reads the positions and bond slots once for the geometry."""
    soup, p = sys._R.soup, sys._R.p
    t = loadStageThresholds()
    box = list(sys.box)
    ctx = ctx or AnalysisContext()
    ctx.load_system(sys)
    particles = sys.particles()
    monomers = soup["monomers"]
    amphiphiles = findAmphiphilesLoaded(ctx, monomers)
    member_idx = memberIndicesOf(amphiphiles)
    kinds = _js_round(particles.reshape(-1, 4)[:, 3].astype(np.float64)).astype(np.int64)
    carbon_kinds = [k for k, m in enumerate(monomers) if m["kind"] == "carbon"]
    total_carbon = int(np.isin(kinds, carbon_kinds).sum())
    carbon_in = sum(a["length"] for a in amphiphiles)
    amph_fraction = carbon_in / total_carbon if total_carbon > 0 else 0
    cutoff = memberCutoff(soup, p)
    largest = 0.0
    enclosed = 0.0
    if member_idx:
        labels = ctx.clusters(np.asarray(member_idx, dtype=np.uint32), box, cutoff, True)
        largest = float(np.bincount(labels).max()) / len(member_idx)
        enclosed = enclosedVolumeFromPositions(ctx, particles, member_idx, box, t["closureCell"], t["closureRadius"])
    head_peaks = computeHeadPeaks(particles, box, monomers, t["headDensityBins"], t["minHeadsPerBin"],
                                  t["headPeakSeparationMin"], t["headPeakSeparationMax"])
    water_id = soup["solvent"]["waterId"]
    water_kind = next((k for k, m in enumerate(monomers) if m["id"] == water_id), -1)
    has_water = water_kind >= 0 and bool((kinds == water_kind).any())
    analysis = analyzeAggregates(ctx, particles, box, monomers, amphiphiles, member_idx, cutoff, t,
                                 water_kind if has_water else None)
    evidence = {"amphiphileFraction": amph_fraction, "largestAggregateFraction": largest, "headPeaks": head_peaks,
                "enclosedVolume": enclosed, "aggregateAnalysis": analysis}
    return stageFromEvidence(evidence, t), evidence

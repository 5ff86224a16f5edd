"""Oracle of the snapshot analysis: soup/src/stages.ts:204-289 (detectStage) restated over the C
primitives of analysis_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of protocell_soup.analysis (GPU graph
and grid work plus the host geometry). The composition below follows soup/src/aggregates.ts and
engine/src/closure.ts statement by statement with plain loops, independently of the product's
vectorised host code; Float32Array stores are float32 stores, JavaScript arithmetic is float64.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from oracle.oracle import lib

NONE = 0xFFFFFFFF
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "protocell-genesis_amd", "data")


def _json(name):
    with open(os.path.join(DATA, name), encoding="utf-8") as fh:
        return json.load(fh)


def _p(a, dtype):
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr, arr.ctypes.data_as(C.c_void_p)


def js_round(x: float) -> float:
    r = math.floor(x)
    return r + 1.0 if x - r >= 0.5 else r


def mi1(d: float, L: float) -> float:
    return d - js_round(d / L) * L


def wrap1(v: float, L: float) -> float:
    return v - math.floor(v / L) * L


def f32(x: float) -> float:
    return float(np.float32(x))


# ---- C primitives -------------------------------------------------------------------------------
def find_amphiphiles(particles, bonds, monomers) -> List[Dict[str, Any]]:
    pos, pp = _p(particles, np.float32)
    n = pos.size // 4
    b, bp = _p(bonds, np.uint32)
    polar, polp = _p([1 if m.get("polar") else 0 for m in monomers], np.uint8)
    head = np.zeros(max(1, n), np.uint32)
    off = np.zeros(n + 1, np.uint32)
    chain = np.zeros(max(1, n), np.uint32)
    tails = np.zeros(max(2, 2 * n), np.uint32)
    na = lib().or_find_amphiphiles(C.c_uint32(n), pp, bp, C.c_uint32(b.size // 2), polp, C.c_uint32(polar.size),
                                   head.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                                   chain.ctypes.data_as(C.c_void_p), tails.ctypes.data_as(C.c_void_p))
    out = []
    for a in range(na):
        ch = [int(x) for x in chain[off[a]:off[a + 1]]]
        tl = [int(tails[2 * a])] + ([int(tails[2 * a + 1])] if tails[2 * a + 1] else [])
        out.append({"headIndex": int(head[a]), "chain": ch, "length": len(ch), "tailLengths": tl})
    return out


def cluster_components(positions, box, cutoff, periodic_z=True) -> List[List[int]]:
    pos, pp = _p(positions, np.float32)
    m = pos.size // 4
    if m == 0:
        return []
    bx = (C.c_double * 3)(*box)
    rank = np.zeros(m, np.uint32)
    size = np.zeros(m, np.uint32)
    nc = lib().or_cluster_components(C.c_uint32(m), pp, bx, C.c_double(cutoff), C.c_int(1 if periodic_z else 0),
                                     rank.ctypes.data_as(C.c_void_p), size.ctypes.data_as(C.c_void_p))
    comps = [[] for _ in range(nc)]
    for k in range(m):
        comps[rank[k]].append(k)
    return comps


def periodic_centre(particles, idx, box, alpha) -> Optional[List[float]]:
    pos, pp = _p(particles, np.float32)
    ix, ip = _p(idx, np.uint32)
    out = (C.c_double * 3)()
    ok = lib().or_periodic_centre(pp, ip, C.c_uint32(ix.size), (C.c_double * 3)(*box), C.c_double(alpha), out)
    return [out[0], out[1], out[2]] if ok else None


def encapsulated_water(particles, member_idx, water_idx, box, centre, cell, radius, enclosed_min) -> Dict[str, Any]:
    pos, pp = _p(particles, np.float32)
    mi, mp = _p(member_idx, np.uint32)
    wi, wp = _p(water_idx, np.uint32)
    out = (C.c_double * 10)()
    lib().or_encapsulated_water(pp, mp, C.c_uint32(mi.size), wp, C.c_uint32(wi.size), (C.c_double * 3)(*box),
                                (C.c_double * 3)(*centre), C.c_double(cell), C.c_double(radius), out)
    if out[6] < 0:
        raise ValueError("farthestEmptyCellSeed: the grid is fully occupied")
    enc, total, reached_w, total_empty, unreached, reached_cells = (out[k] for k in range(6))
    bulk = reached_w / (reached_cells * cell ** 3) if reached_cells > 0 else 0.0
    thr = enclosed_min * bulk if bulk > 0 else float("nan")
    seed, nx, ny = int(out[6]), int(out[7]), int(out[8])
    z = seed // (nx * ny)
    rem = seed - z * nx * ny
    y = rem // nx
    x = rem - y * nx
    dx = mi1((x + 0.5) * cell - centre[0], box[0])
    dy = mi1((y + 0.5) * cell - centre[1], box[1])
    dz = mi1((z + 0.5) * cell - centre[2], box[2])
    return {"encapsulatedCount": int(enc), "totalWater": int(total), "bulkWaterDensity": bulk,
            "encapsulationThresholdCount": thr, "closed": bool(math.isfinite(thr) and enc > thr),
            "seedDistFromCentre": math.sqrt(dx * dx + dy * dy + dz * dz),
            "theoreticalMaxDist": (math.sqrt(3) * min(box)) / 2, "totalEmptyCells": int(total_empty),
            "unreachedCells": int(unreached)}


def enclosed_cells(positions, box, cell, radius) -> int:
    pos, pp = _p(positions, np.float32)
    lib().or_enclosed_cells.restype = C.c_double
    return int(lib().or_enclosed_cells(pp, C.c_uint32(pos.size // 4), (C.c_double * 3)(*box), C.c_double(cell),
                                       C.c_double(radius)))


def shape_of(positions) -> Dict[str, Any]:
    pos, pp = _p(positions, np.float32)
    out = (C.c_double * 10)()
    lib().or_shape_of(pp, C.c_uint32(pos.size // 4), out)
    return {"centre": [out[0], out[1], out[2]], "radiusOfGyration": out[3], "principalMoments": [out[4], out[5], out[6]],
            "smallestAxis": [out[7], out[8], out[9]]}


def head_peaks(particles, box, monomers, bins, min_per_bin, sep_min, sep_max):
    pos, pp = _p(particles, np.float32)
    polar, polp = _p([1 if m.get("polar") else 0 for m in monomers], np.uint8)
    sigma = _json("params.json")["sigma"]
    r = lib().or_head_peaks(pp, C.c_uint32(pos.size // 4), (C.c_double * 3)(*box), polp, C.c_uint32(polar.size),
                            C.c_uint32(bins), C.c_double(min_per_bin), C.c_double(sep_min), C.c_double(sep_max),
                            C.c_double(sigma))
    return "unavailable" if r < 0 else int(r)


# ---- composition (soup/src/aggregates.ts, soup/src/stages.ts) -----------------------------------
def member_indices(amphiphiles) -> List[int]:
    out, seen = [], set()
    for a in amphiphiles:
        for i in [a["headIndex"]] + a["chain"]:
            if i not in seen:
                seen.add(i)
                out.append(i)
    return out


def _unwrap(raw: np.ndarray, box) -> np.ndarray:
    out = np.zeros_like(raw)
    n = raw.shape[0]
    ref = [float(raw[0, a]) for a in range(3)] if n else [0.0, 0.0, 0.0]
    for i in range(n):
        for a in range(3):
            out[i, a] = f32(ref[a] + mi1(float(raw[i, a]) - ref[a], box[a]))
        out[i, 3] = raw[i, 3]
    return out


def _profile(coords: List[float], kinds) -> tuple:
    lo, hi = min(coords), max(coords)
    span = max(1e-6, (hi - lo) * 1.0001)
    p = np.zeros((len(coords), 4), np.float32)
    for i, c in enumerate(coords):
        p[i, 2] = f32(c - lo)
        p[i, 3] = kinds[i]
    return p, [1.0, 1.0, span]


def _local_cavity(unwrapped: np.ndarray, cell, radius) -> float:
    n = unwrapped.shape[0]
    if n == 0:
        return 0.0
    lo = [min(float(v) for v in unwrapped[:, a]) for a in range(3)]
    hi = [max(float(v) for v in unwrapped[:, a]) for a in range(3)]
    pad = radius * 4
    shift = [pad - lo[a] for a in range(3)]
    lbox = [hi[a] - lo[a] + 2 * pad for a in range(3)]
    sh = unwrapped.copy()
    for i in range(n):
        for a in range(3):
            sh[i, a] = f32(float(unwrapped[i, a]) + shift[a])
    return enclosed_cells(sh, lbox, cell, radius) * cell * cell * cell


def analyze_aggregates(particles, box, monomers, amphiphiles, member_idx, cutoff, t, water_idx=None) -> Dict[str, Any]:
    pos = np.asarray(particles, dtype=np.float32).reshape(-1, 4)
    if not member_idx:
        return {"aggregateCount": 0, "sizeHistogram": [], "aggregates": [], "qualifyingAggregateCount": 0,
                "amphiphilesInQualifying": 0, "amphiphileShareInQualifying": 0, "hasLamellarAggregate": False,
                "hasVesicleAggregate": False}
    comps = cluster_components(pos[member_idx], box, cutoff, True)
    heads = {a["headIndex"] for a in amphiphiles}
    prepared = []
    for local in comps:
        orig = [member_idx[k] for k in local]
        prepared.append((orig, sum(1 for i in orig if i in heads)))
    q = [c for _, c in prepared if c >= t["minAmphiphilesPerAggregate"]]
    aggs = []
    alpha = _json("periodic-measurement.json")["circularConcentration"]["alpha"]
    for orig, count in prepared[: min(t["detailAggregateCount"], len(prepared))]:
        un = _unwrap(pos[orig], box)
        shp = shape_of(un)
        l0, l1, l2 = shp["principalMoments"]
        flat = l0 / l2 if l2 > 1e-12 else 1
        sym = l1 / l2 if l2 > 1e-12 else 1
        ctr = shp["centre"]
        kinds = [float(k) for k in un[:, 3]]
        radial = [math.sqrt((float(un[i, 0]) - ctr[0]) ** 2 + (float(un[i, 1]) - ctr[1]) ** 2 + (float(un[i, 2]) - ctr[2]) ** 2)
                  for i in range(un.shape[0])]
        hp = (t["aggregateHeadDensityBins"], t["aggregateMinHeadsPerBin"], t["headPeakSeparationMin"], t["headPeakSeparationMax"])
        rp, rb = _profile(radial, kinds)
        rsh = head_peaks(rp, rb, monomers, *hp)
        tsh = 0
        if flat <= t["lamellarFlatnessRatio"] and sym >= t["lamellarInPlaneSymmetryMin"]:
            ax = shp["smallestAxis"]
            tr = [(float(un[i, 0]) - ctr[0]) * ax[0] + (float(un[i, 1]) - ctr[1]) * ax[1] + (float(un[i, 2]) - ctr[2]) * ax[2]
                  for i in range(un.shape[0])]
            tp, tb = _profile(tr, kinds)
            tsh = head_peaks(tp, tb, monomers, *hp)
        s = {"particleCount": un.shape[0], "amphiphileCount": count, "centre": ctr,
             "radiusOfGyration": shp["radiusOfGyration"], "principalMoments": shp["principalMoments"],
             "flatnessRatio": flat, "inPlaneSymmetry": sym, "radialHeadShells": rsh, "transverseHeadShells": tsh,
             "cavityVolume": _local_cavity(un, t["closureCell"], t["closureRadius"])}
        if water_idx:
            c = periodic_centre(pos, orig, box, alpha)
            s["encapsulatedWater"] = (encapsulated_water(pos, orig, water_idx, box, c, t["closureCell"], t["closureRadius"],
                                                         t["enclosedVolume"]) if c is not None else None)
        aggs.append(s)

    def lamellar(s):
        return s["flatnessRatio"] <= t["lamellarFlatnessRatio"] and s["inPlaneSymmetry"] >= t["lamellarInPlaneSymmetryMin"]

    def vesicle(s):
        if s["radialHeadShells"] != 2:
            return False
        if "encapsulatedWater" in s:
            return s["encapsulatedWater"] is not None and s["encapsulatedWater"]["closed"]
        return s["cavityVolume"] > t["enclosedVolume"]

    return {"aggregateCount": len(comps), "sizeHistogram": [c for _, c in prepared], "aggregates": aggs,
            "qualifyingAggregateCount": len(q), "amphiphilesInQualifying": sum(q),
            "amphiphileShareInQualifying": sum(q) / len(amphiphiles) if amphiphiles else 0,
            "hasLamellarAggregate": any(lamellar(s) and (s["radialHeadShells"] == 2 or s["transverseHeadShells"] == 2)
                                        for s in aggs),
            "hasVesicleAggregate": any(vesicle(s) for s in aggs)}


def enclosed_volume_box(particles, member_idx, box, cell, radius) -> float:
    """enclosedVolumeFromPositions (engine/src/closure.ts:291-323) over the member beads."""
    params = _json("params.json")
    cut = math.pow(2.0, 1.0 / 6.0) * params["beadSizes"]["tail_tail"] + params["attraction"]["wc"]
    pos = np.asarray(particles, dtype=np.float32).reshape(-1, 4)[member_idx]
    comps = cluster_components(pos, box, cut, False)
    # the first component (in first-occurrence order) of the largest size: clusterComponents sorts
    # stably by size, so that is comps[0]
    best = comps[0]
    ref_x, ref_y = float(pos[best[0], 0]), float(pos[best[0], 1])
    sx = sy = sz = 0.0
    for k in best:
        sx += mi1(float(pos[k, 0]) - ref_x, box[0])
        sy += mi1(float(pos[k, 1]) - ref_y, box[1])
        sz += float(pos[k, 2])
    n = len(best)
    com = [wrap1(ref_x + sx / n, box[0]), wrap1(ref_y + sy / n, box[1]), sz / n]
    shift = [box[0] / 2 - com[0], box[1] / 2 - com[1], box[2] / 2 - com[2]]
    out = pos.copy()
    for i in range(pos.shape[0]):
        out[i, 0] = f32(wrap1(float(pos[i, 0]) + shift[0], box[0]))
        out[i, 1] = f32(wrap1(float(pos[i, 1]) + shift[1], box[1]))
        out[i, 2] = f32(float(pos[i, 2]) + shift[2])
    return enclosed_cells(out, box, cell, radius) * cell * cell * cell

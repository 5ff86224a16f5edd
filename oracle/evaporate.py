"""Oracle of the solvent insertion and the returning beads' state: soup/src/soup-evaporate-plan.ts:202-304
(sampleInsertionPositions) and soup/src/soup-evaporate.ts:166-212 (the mulberry32 stream, the
gaussian velocity draws and the per-index RNG seeds), restated with plain Python loops.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's soup_sample_insertion
and of protocell_soup.evaporate. JavaScript numbers are float64, Float32Array stores are float32
stores, Math.imul and >>> are 32-bit integer operations.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence

import numpy as np

INSERTION_CANDIDATES = 64  # soup-evaporate-plan.ts:25


def _imul(a: int, b: int) -> int:
    return (a * b) & 0xFFFFFFFF


class Mulberry32:
    """The closure of soup-evaporate.ts:169-174 over its 32-bit state ``a``."""

    def __init__(self, a: int):
        self.a = a & 0xFFFFFFFF

    def __call__(self) -> float:
        self.a = (self.a + 0x6D2B79F5) & 0xFFFFFFFF
        a = self.a
        t = _imul(a ^ (a >> 15), 1 | a)
        t = ((t + _imul(t ^ (t >> 7), 61 | t)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0


def rehydration_seed(seed: int, global_step: int) -> int:
    """soup-evaporate.ts:169: (seed ^ Math.imul(globalStep + 1, 0x9e3779b9)) >>> 0."""
    return ((seed & 0xFFFFFFFF) ^ _imul((global_step + 1) & 0xFFFFFFFF, 0x9E3779B9)) & 0xFFFFFFFF


def js_round(x: float) -> float:
    """Math.round: the nearest integer, halves toward +infinity."""
    r = math.floor(x)
    return float(r + 1) if x - r >= 0.5 else float(r)


def sample_insertion_positions(active: np.ndarray, active_n: int, box: Sequence[float], count: int, min_sep: float,
                               rng: Mulberry32) -> Dict:
    """soup-evaporate-plan.ts:202-304 statement by statement (no parcel draw)."""
    nx = max(1, math.floor(box[0] / min_sep))
    ny = max(1, math.floor(box[1] / min_sep))
    nz = max(1, math.floor(box[2] / min_sep))
    wx, wy, wz = box[0] / nx, box[1] / ny, box[2] / nz
    cells: Dict[int, List[int]] = {}
    xs: List[float] = []
    ys: List[float] = []
    zs: List[float] = []

    def cell_of(x, y, z):
        cx = min(nx - 1, math.floor(x / wx))
        cy = min(ny - 1, math.floor(y / wy))
        cz = min(nz - 1, math.floor(z / wz))
        return cx + nx * (cy + ny * cz)

    def push(x, y, z):
        idx = len(xs)
        xs.append(x)
        ys.append(y)
        zs.append(z)
        cells.setdefault(cell_of(x, y, z), []).append(idx)

    act = np.asarray(active, dtype=np.float32).reshape(-1)
    for i in range(active_n):
        push(float(act[i * 4]), float(act[i * 4 + 1]), float(act[i * 4 + 2]))

    def nearest(x, y, z):
        cx = min(nx - 1, math.floor(x / wx))
        cy = min(ny - 1, math.floor(y / wy))
        cz = min(nz - 1, math.floor(z / wz))
        best = math.inf
        for dx in (-1, 0, 1):
            ax = ((cx + dx) % nx + nx) % nx
            for dy in (-1, 0, 1):
                ay = ((cy + dy) % ny + ny) % ny
                for dz in (-1, 0, 1):
                    az = ((cz + dz) % nz + nz) % nz
                    for j in cells.get(ax + nx * (ay + ny * az), ()):
                        d2 = 0.0
                        for a, bb, L in ((x, xs[j], box[0]), (y, ys[j], box[1]), (z, zs[j], box[2])):
                            d = a - bb
                            m = d - js_round(d / L) * L
                            d2 += m * m
                        if d2 < best:
                            best = d2
        return math.sqrt(best)

    positions = np.zeros(count * 4, dtype=np.float32)
    short = 0
    min_achieved = math.inf
    for k in range(count):
        bx = by = bz = 0.0
        bd = -1.0
        for _ in range(INSERTION_CANDIDATES):
            x = rng() * box[0]
            y = rng() * box[1]
            z = rng() * box[2]
            d = nearest(x, y, z)
            if d > bd:
                bd, bx, by, bz = d, x, y, z
            if d >= min_sep:
                break
        if bd < min_sep:
            short += 1
        if bd < min_achieved:
            min_achieved = bd
        positions[k * 4] = bx
        positions[k * 4 + 1] = by
        positions[k * 4 + 2] = bz
        push(bx, by, bz)
    return {"positions": positions, "shortOfFloor": short,
            "minAchieved": min_achieved if math.isfinite(min_achieved) else 0.0}


def returning_beads(seed: int, global_step: int, active: np.ndarray, insert_start: int, box: Sequence[float], count: int,
                    min_sep: float, kT: float, solvent_kind: int) -> Dict[str, np.ndarray]:
    """The arrays rehydrateSolventTo writes for its `count` beads (soup-evaporate.ts:166-212):
    positions from the sampler, velocities s * gaussian() continuing the same stream, and the two
    per-index RNG seeds."""
    rng = Mulberry32(rehydration_seed(seed, global_step))
    sampled = sample_insertion_positions(active, insert_start, box, count, min_sep, rng)

    def gaussian():
        u1 = max(rng(), 1e-9)
        u2 = rng()
        return math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2)

    pos = np.zeros(count * 4, dtype=np.float32)
    vel = np.zeros(count * 4, dtype=np.float32)
    s = math.sqrt(kT)
    for k in range(count):
        pos[k * 4:k * 4 + 3] = sampled["positions"][k * 4:k * 4 + 3]
        pos[k * 4 + 3] = solvent_kind
        vel[k * 4] = s * gaussian()
        vel[k * 4 + 1] = s * gaussian()
        vel[k * 4 + 2] = s * gaussian()
    bond = np.zeros(count, dtype=np.uint32)
    thermo = np.zeros(count, dtype=np.uint32)
    for k in range(count):
        i = insert_start + k
        sd = seed & 0xFFFFFFFF
        g = global_step & 0xFFFFFFFF
        bond[k] = sd ^ _imul(i + 1, 2654435761) ^ 0x9E3779B9 ^ g
        thermo[k] = sd ^ _imul(i + 1, 3266489917) ^ 0x85EBCA6B ^ g
    return {"positions": pos, "velocities": vel, "bondRng": bond, "thermoRng": thermo,
            "shortOfFloor": sampled["shortOfFloor"], "minAchieved": sampled["minAchieved"]}

"""Initial state: the seeded jittered lattice and the per-particle RNG seeds.

Restates soup/src/soup-init-state.ts:17-31 (mulberry32, gaussian) and :58-307
(``buildInitialState``: Fisher-Yates site shuffle, 0.15 jitter, Maxwell velocities, per-particle
bond/thermostat seeds at :269-272) for the clay-free, unconfined case. mulberry32's state advances
by a constant, so the k-th draw is a pure function of k and the whole draw sequence is generated
vectorised; only the Fisher-Yates swaps run as a scalar loop.

Also builds pre-assembled amphiphile systems by writing ``bondSlots`` the way
tests/water-bilayer-area-move.test.ts:131-134 does (head slot 0 -> tail1, tail1 slot 2 -> head,
forward links in slot 0, backward links between tails in slot 1, the terminal tail's only link in
slot 0).
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence

import numpy as np

NONE_U32 = 0xFFFFFFFF
_GOLDEN = 0x6D2B79F5


def mulberry32_draws(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Draws start..start+count-1 of mulberry32(seed) as float64 in [0, 1)."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    a = ((np.uint64(seed & 0xFFFFFFFF) + k * np.uint64(_GOLDEN)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    t = (a ^ (a >> np.uint32(15))) * (a | np.uint32(1))
    t = (t + (t ^ (t >> np.uint32(7))) * (t | np.uint32(61))) ^ t
    return (t ^ (t >> np.uint32(14))).astype(np.float64) / 4294967296.0


class Mulberry32:
    """Scalar mulberry32 cursor over the vectorised draw sequence."""

    def __init__(self, seed: int):
        self.seed = seed & 0xFFFFFFFF
        self.k = 0

    def take(self, count: int) -> np.ndarray:
        out = mulberry32_draws(self.seed, count, self.k)
        self.k += count
        return out


def _bm(u1: np.ndarray, u2: np.ndarray) -> np.ndarray:
    """gaussian() (soup-init-state.ts:27-31) on pre-drawn pairs."""
    return np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-9))) * np.cos(2.0 * math.pi * u2)


def rng_seeds(seed: int, n: int):
    """bondRng0 / thermoRng0 (soup-init-state.ts:269-272)."""
    i1 = np.arange(1, n + 1, dtype=np.uint64)
    s = np.uint64(seed & 0xFFFFFFFF)
    bond = (s ^ ((i1 * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) ^ np.uint64(0x9E3779B9)).astype(np.uint32)
    thermo = (s ^ ((i1 * np.uint64(3266489917)) & np.uint64(0xFFFFFFFF)) ^ np.uint64(0x85EBCA6B)).astype(np.uint32)
    return bond, thermo


def lattice_state(box: Sequence[float], counts_by_kind: Sequence[int], seed: int, kT: float) -> Dict[str, np.ndarray]:
    """Fresh-creation branch of buildInitialState (no clay, no parcel, no frozen catalysts)."""
    n = int(sum(counts_by_kind))
    rng = Mulberry32(seed)
    # Math.max(1, Math.ceil(Math.cbrt(N))) then up while nx^3 < N (soup-init-state.ts:117-118): V8's
    # cbrt is exact on perfect cubes, so this is the smallest nx with nx^3 >= N, found in integers
    # (a libm cbrt one ulp above an integer would give nx + 1 for a perfect cube)
    nx = 1
    while nx * nx * nx < n:
        nx += 1
    spacing = np.array([box[0] / nx, box[1] / nx, box[2] / nx], dtype=np.float64)
    n_sites = nx * nx * nx
    order = np.arange(n_sites, dtype=np.int64)
    if n_sites > 1:
        draws = rng.take(n_sites - 1)
        ii = np.arange(n_sites - 1, 0, -1, dtype=np.int64)
        js = np.floor(draws * (ii + 1)).astype(np.int64)
        ol = order.tolist()
        for i, j in zip(ii.tolist(), js.tolist()):
            ol[i], ol[j] = ol[j], ol[i]
        order = np.array(ol, dtype=np.int64)
    d = rng.take(9 * n).reshape(n, 9)
    site = order[:n]
    ix, iy, iz = site % nx, (site // nx) % nx, site // (nx * nx)
    jf = 0.15
    x = (ix + 0.5) * spacing[0] + (d[:, 0] * 2 - 1) * jf * spacing[0]
    y = (iy + 0.5) * spacing[1] + (d[:, 1] * 2 - 1) * jf * spacing[1]
    z = (iz + 0.5) * spacing[2] + (d[:, 2] * 2 - 1) * jf * spacing[2]
    kinds = np.repeat(np.arange(len(counts_by_kind)), counts_by_kind).astype(np.float64)
    pos = np.stack([x, y, z, kinds], axis=1).astype(np.float32)
    s = math.sqrt(kT)
    vel = np.zeros((n, 4), dtype=np.float32)
    vel[:, 0] = (s * _bm(d[:, 3], d[:, 4])).astype(np.float32)
    vel[:, 1] = (s * _bm(d[:, 5], d[:, 6])).astype(np.float32)
    vel[:, 2] = (s * _bm(d[:, 7], d[:, 8])).astype(np.float32)
    bond, thermo = rng_seeds(seed, n)
    return {
        "positions": pos.reshape(-1),
        "velocities": vel.reshape(-1),
        "bondSlots": np.full(3 * n, NONE_U32, dtype=np.uint32),
        "centerLink": np.full(n, NONE_U32, dtype=np.uint32),
        "centerHeldSteps": np.zeros(n, dtype=np.uint32),
        "frozen": np.zeros(n, dtype=np.uint32),
        "desorbEvents": np.zeros(2, dtype=np.uint32),
        "bondRng": bond,
        "thermoRng": thermo,
    }


def link_chain(bond_slots: np.ndarray, head: int, tails: Sequence[int]) -> None:
    """Writes one linear amphiphile head-t1-...-tk into bondSlots (3 per particle)."""
    bs = bond_slots
    if len(tails) == 0:
        return
    bs[3 * head + 0] = tails[0]
    bs[3 * tails[0] + 2] = head
    for a in range(len(tails) - 1):
        t, u = tails[a], tails[a + 1]
        bs[3 * t + 0] = u
        back_slot = 1 if a + 1 < len(tails) - 1 else 0
        bs[3 * u + back_slot] = t


def amphiphile_soup(box: Sequence[float], n_amph: int, tail_len: int, extra_counts: Dict[int, int], head_kind: int,
                    carbon_kind: int, n_kinds: int, seed: int, kT: float) -> Dict[str, np.ndarray]:
    """A jittered cubic lattice with ``n_amph`` pre-bonded linear amphiphiles (1 head + tail_len
    carbons on consecutive x-sites of one lattice row) and free beads of the other kinds on the
    remaining sites. Particle order: all amphiphile beads first (head, t1..tk per molecule), then the
    free beads in kind order. The kind census is what createSoup's ``start`` must carry."""
    per = 1 + tail_len
    n_free = int(sum(extra_counts.values()))
    n = n_amph * per + n_free
    rng = Mulberry32(seed)
    # Math.max(1, Math.ceil(Math.cbrt(N))) then up while nx^3 < N (soup-init-state.ts:117-118): V8's
    # cbrt is exact on perfect cubes, so this is the smallest nx with nx^3 >= N, found in integers
    # (a libm cbrt one ulp above an integer would give nx + 1 for a perfect cube)
    nx = 1
    while nx * nx * nx < n:
        nx += 1
    seg_per_row = nx // per
    if seg_per_row == 0:
        raise ValueError("amphiphile_soup: box too small for one molecule per lattice row")
    n_rows = nx * nx
    n_seg = n_rows * seg_per_row
    if n_amph > n_seg:
        raise ValueError("amphiphile_soup: not enough lattice rows for the requested amphiphiles")
    # choose segments by a seeded shuffle
    seg_order = np.argsort(rng.take(n_seg), kind="stable")[:n_amph]
    used = np.zeros(nx * nx * nx, dtype=bool)
    site_of = np.empty(n, dtype=np.int64)
    kinds = np.empty(n, dtype=np.float32)
    k = 0
    bs = np.full(3 * n, NONE_U32, dtype=np.uint32)
    for s in seg_order.tolist():
        row, slot = divmod(s, seg_per_row)
        iy, iz = row % nx, row // nx
        x0 = slot * per
        idx = []
        for b in range(per):
            site = (x0 + b) + nx * (iy + nx * iz)
            used[site] = True
            site_of[k] = site
            kinds[k] = head_kind if b == 0 else carbon_kind
            idx.append(k)
            k += 1
        link_chain(bs, idx[0], idx[1:])
    free_sites = np.nonzero(~used)[0]
    free_sites = free_sites[np.argsort(rng.take(free_sites.size), kind="stable")]
    f = 0
    for kind in range(n_kinds):
        for _ in range(int(extra_counts.get(kind, 0))):
            site_of[k] = free_sites[f]
            kinds[k] = kind
            k += 1
            f += 1
    spacing = np.array(box, dtype=np.float64) / nx
    d = rng.take(9 * n).reshape(n, 9)
    ix, iy, iz = site_of % nx, (site_of // nx) % nx, site_of // (nx * nx)
    jf = 0.15
    pos = np.empty((n, 4), dtype=np.float32)
    for a, ia in enumerate((ix, iy, iz)):
        pos[:, a] = ((ia + 0.5) * spacing[a] + (d[:, a] * 2 - 1) * jf * spacing[a]).astype(np.float32)
    # chain beads sit on a straight rod at a bond length below the FENE divergence radius, starting
    # at the segment's first site, with the jitter kept perpendicular to the rod
    bond = min(float(spacing[0]), 1.0)
    n_chain = n_amph * per
    b = np.arange(n_chain) % per
    x0 = ix[:n_chain] - b
    pos[:n_chain, 0] = ((x0 + 0.5) * spacing[0] + b * bond).astype(np.float32)
    for a, ia in ((1, iy), (2, iz)):
        pos[:n_chain, a] = ((ia[:n_chain] + 0.5) * spacing[a] + (d[:n_chain, a] * 2 - 1) * 0.5 * jf * spacing[a]).astype(np.float32)
    pos[:, :3] = np.mod(pos[:, :3], np.array(box, dtype=np.float32))
    pos[:, 3] = kinds
    s = math.sqrt(kT)
    vel = np.zeros((n, 4), dtype=np.float32)
    for a in range(3):
        vel[:, a] = (s * _bm(d[:, 3 + 2 * a], d[:, 4 + 2 * a])).astype(np.float32)
    bond, thermo = rng_seeds(seed, n)
    return {
        "positions": pos.reshape(-1),
        "velocities": vel.reshape(-1),
        "bondSlots": bs,
        "centerLink": np.full(n, NONE_U32, dtype=np.uint32),
        "centerHeldSteps": np.zeros(n, dtype=np.uint32),
        "frozen": np.zeros(n, dtype=np.uint32),
        "desorbEvents": np.zeros(2, dtype=np.uint32),
        "bondRng": bond,
        "thermoRng": thermo,
    }

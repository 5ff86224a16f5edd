"""The zero-tension Metropolis area move (soup/src/soup-area-move.ts).

The chain runs on the host as the reference's does, one proposal at a time:
- each trial box is mapped molecule by molecule (scaleMoleculesRigid, native host code);
- its FP64 soupPotential is evaluated on the GPU from the configuration's own cell grid
  (soup_potential_at);
- the Metropolis test uses the reference's mulberry32 stream and its Jacobian term.

The accepted chain is applied to the device once, at the box it ended on, through
box_scale.applyBoxScaleOnce: the rigid map, the reference's bonded-distance self-check, then
soup_set_box, which adopts the new grid and runs a rebuild and a force pass.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np

from . import native
from .init_state import Mulberry32
from .box_scale import applyBoxScaleOnce  # noqa: F401  (the box change an accepted chain applies)
from .plan import planSoupGrid

Box = Tuple[float, float, float]


def proposeBox(box: Sequence[float], u: float, mode: str) -> Box:
    """soup-area-move.ts:60-64: the lateral axes scale by sqrt(exp(u)); 'lateral-fixed-volume'
    divides L_z by exp(u), 'lateral-fixed-z' keeps it."""
    s = math.exp(u)
    q = math.sqrt(s)
    if mode == "lateral-fixed-volume":
        return (box[0] * q, box[1] * q, box[2] / s)
    return (box[0] * q, box[1] * q, box[2])


def scaleBeadsAffine(positions: np.ndarray, oldBox: Sequence[float], newBox: Sequence[float]) -> np.ndarray:
    """soup-area-move.ts:72-84: every bead rescaled on its own (the verification map only)."""
    p = np.asarray(positions, np.float32).reshape(-1, 4)
    out = np.empty_like(p)
    for a in range(3):
        v = p[:, a].astype(np.float64) * (newBox[a] / oldBox[a])
        out[:, a] = (v - np.floor(v / newBox[a]) * newBox[a]).astype(np.float32)
    out[:, 3] = p[:, 3]
    return out.reshape(-1)


def moleculeCount(bonds: np.ndarray, n: int) -> int:
    """soup-area-move.ts:91-115: connected components of the bond graph, every unbonded particle
    its own component (the N of the configurational Jacobian)."""
    parent = np.arange(n, dtype=np.int64)

    def find(x: int) -> int:
        r = x
        while parent[r] != r:
            r = int(parent[r])
        while parent[x] != r:
            parent[x], x = r, int(parent[x])
        return r

    comps = n
    b = np.asarray(bonds, np.int64).reshape(-1)
    for k in range(0, b.size, 2):
        ra, rb = find(int(b[k])), find(int(b[k + 1]))
        if ra != rb:
            parent[ra] = rb
            comps -= 1
    return comps


def scaleMoleculesRigid(positions: np.ndarray, bonds: np.ndarray, oldBox: Sequence[float], newBox: Sequence[float]) -> np.ndarray:
    """soup-box-scale-math.ts:63-138 through the library's host implementation
    (soup_scale_molecules_rigid)."""
    lib = native.load_library()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    n = pos.size // 4
    pairs = np.ascontiguousarray(bonds, dtype=np.uint32).reshape(-1)
    out = np.empty_like(pos)
    ob = (C.c_double * 3)(*[float(x) for x in oldBox])
    nb = (C.c_double * 3)(*[float(x) for x in newBox])
    r = lib.soup_scale_molecules_rigid(pos.ctypes.data_as(C.c_void_p), C.c_uint32(n), pairs.ctypes.data_as(C.c_void_p),
                                       C.c_uint32(pairs.size // 2), ob, nb, out.ctypes.data_as(C.c_void_p))
    if r != 0:
        raise ValueError("scaleMoleculesRigid: a bond names a particle outside the system")
    return out


class SoupAreaMove:
    """makeSoupAreaMove (soup-area-move.ts:167-279) bound to one system: one mulberry32 stream
    (seed ^ 0x5eeda1ea) for the system's lifetime, the potential basis built on first use."""

    def __init__(self, sys_):
        self._sys = sys_
        self._rng = Mulberry32((int(sys_._R.seed) ^ 0x5EEDA1EA) & 0xFFFFFFFF)
        self._basis = None

    def _energy(self, positions: np.ndarray, box: Sequence[float]) -> float:
        s = self._sys
        out = (C.c_double * 5)()
        bx = (C.c_double * 3)(*[float(x) for x in box])
        pos = np.ascontiguousarray(positions, dtype=np.float32)
        s._call(s._lib.soup_potential_at, C.byref(self._basis), pos.ctypes.data_as(C.c_void_p), bx, out)
        return float(out[4])

    def __call__(self, trials: int, opts: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
        s = self._sys
        R = s._R
        am = R.soup.get("areaMove")
        if am is None:
            raise ValueError("data/soup.json: no areaMove section, so the MC area move is not configured (see soup/src/soup-area-move.ts)")
        mode = (opts or {}).get("mode", am["mode"])
        if self._basis is None:
            self._basis = s.potentialBasis(centerLink=True, charges=True)
        start_box = tuple(float(x) for x in R.live_box)
        pairs = s.bonds()
        cur_pos = s.particles()
        cur_box = start_box
        kT = float(R.p["thermostat"]["kT"])
        n_mol = moleculeCount(pairs, R.N)
        energy = self._energy(cur_pos, cur_box)
        energy_start = energy
        trajectory = []
        accepted = 0
        for _ in range(int(trials)):
            u = (float(self._rng.take(1)[0]) * 2.0 - 1.0) * float(am["logDelta"])
            prop_box = proposeBox(cur_box, u, mode)
            if not planSoupGrid(prop_box, R.start).valid:
                trajectory.append(cur_box[0])
                continue
            prop_pos = scaleMoleculesRigid(cur_pos, pairs, cur_box, prop_box)
            prop_energy = self._energy(prop_pos, prop_box)
            dU = prop_energy - energy
            vol_ratio = (prop_box[0] * prop_box[1] * prop_box[2]) / (cur_box[0] * cur_box[1] * cur_box[2])
            accept_prob = min(1.0, math.exp(n_mol * math.log(vol_ratio) - dU / kT))
            if float(self._rng.take(1)[0]) < accept_prob:
                accepted += 1
                cur_pos, cur_box, energy = prop_pos, prop_box, prop_energy
            trajectory.append(cur_box[0])
        if cur_box != start_box:
            applyBoxScaleOnce(s, cur_box)
        return {"trials": int(trials), "accepted": accepted, "acceptedFraction": accepted / trials if trials > 0 else 0.0,
                "box": list(cur_box), "lateralTrajectory": trajectory, "energyStart": energy_start, "energyEnd": energy}

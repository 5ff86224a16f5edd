"""Gradients of the soup energy and forces (soup_energy_adjoint) and their torch.autograd front.

U(x, theta) is soupPotential (soup/src/soup-potential.ts:145-301) over the system's bonds, centre
links and charges. theta collects the parameters the adjoint differentiates, in this order:
- the attraction depth of each unordered class pair, attr(0,0) ... attr(3,3) (10 values; both
  halves of the symmetric table together);
- the acid-soap depth, epsilon (WCA and attraction scale), the head charge (chargeDeprotonated)
  and the DH screening kappa (at fixed coefficient and cutoff; the shifts follow kappa).

SoupEnergy and SoupForce are torch.autograd.Functions over a live system: the positions tensor is
written to the device (rounded to f32, the precision the system holds), the energy or the FP64
force is evaluated there, and backward is one adjoint launch. SoupForce's backward is the
vector-Jacobian product of the force, -H w for the positions and d(w . F)/dtheta for theta, so a
loss on forces (force matching, a virial) differentiates through it.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from . import native

CLASS_PAIRS = [(a, b) for a in range(4) for b in range(a, 4)]
PARAM_NAMES: List[str] = [f"attr{a}{b}" for a, b in CLASS_PAIRS] + ["acidSoap", "epsilon", "headCharge", "kappa"]
assert len(PARAM_NAMES) == 14


def copyBasis(b: native.SoupPotentialBasis) -> native.SoupPotentialBasis:
    out = native.SoupPotentialBasis()
    C.memmove(C.byref(out), C.byref(b), C.sizeof(native.SoupPotentialBasis))
    return out


def paramsOf(b: native.SoupPotentialBasis, headCharge: float) -> np.ndarray:
    """theta of a basis (the head charge is the system's chargeDeprotonated, not a basis field)."""
    th = [b.attr[a][c] for a, c in CLASS_PAIRS] + [b.acid_soap, b.epsilon, headCharge, b.es_kappa]
    return np.asarray(th, dtype=np.float64)


def basisWith(b: native.SoupPotentialBasis, theta: Sequence[float], headCharge: float) -> native.SoupPotentialBasis:
    """A copy of `b` at parameters theta. A head charge q' != q enters as the DH coefficient and shifts
    scaled by (q'/q)^2, which is what every charged pair sees; a new kappa recomputes both shifts."""
    out = copyBasis(b)
    th = np.asarray(theta, dtype=np.float64)
    for k, (a, c) in enumerate(CLASS_PAIRS):
        out.attr[a][c] = float(th[k])
        out.attr[c][a] = float(th[k])
    out.acid_soap = float(th[10])
    out.epsilon = float(th[11])
    scale = (float(th[12]) / headCharge) ** 2 if headCharge != 0.0 else 1.0
    kap = float(th[13])
    A = b.es_A * scale
    rc = b.es_cutoff
    out.es_A = A
    out.es_kappa = kap
    if rc > 0.0:
        out.es_shiftU = A * math.exp(-kap * rc) / rc
        out.es_shiftF = out.es_shiftU * (1.0 / rc + kap)
    return out


def energyAdjoint(sys_, basis: Optional[native.SoupPotentialBasis] = None, w: Optional[np.ndarray] = None) -> Dict[str, Any]:
    """One adjoint launch at the system's configuration: grad = dU/dx (N, 3), dU = dU/dtheta, and with
    a direction w (N, 3): hw = H w and dFw = d(w . F)/dtheta. FP64 throughout."""
    b = basis if basis is not None else sys_.potentialBasis()
    n = sys_.N
    grad = np.empty(3 * n, dtype=np.float64)
    dU = np.empty(14, dtype=np.float64)
    hw = dFw = None
    wp = hp = fp = None
    if w is not None:
        wa = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        if wa.size != 3 * n:
            raise ValueError(f"energyAdjoint: w holds {wa.size // 3} particles, the system {n}")
        hw = np.empty(3 * n, dtype=np.float64)
        dFw = np.empty(14, dtype=np.float64)
        wp, hp, fp = wa.ctypes.data_as(C.c_void_p), hw.ctypes.data_as(C.c_void_p), dFw.ctypes.data_as(C.c_void_p)
    sys_._call(sys_._lib.soup_energy_adjoint, C.byref(b), wp, grad.ctypes.data_as(C.c_void_p), hp,
               dU.ctypes.data_as(C.c_void_p), fp)
    out = {"grad": grad.reshape(n, 3), "dU": dU}
    if w is not None:
        out["hw"] = hw.reshape(n, 3)
        out["dFw"] = dFw
    return out


def setPositions(sys_, x: np.ndarray) -> None:
    """Writes positions (N, 3) into the system, keeping every kind; a move beyond half the skin
    since the last rebuild makes the next list-based call refuse, as the drift guard does."""
    p = sys_.particles().reshape(-1, 4).copy()
    p[:, :3] = np.asarray(x, dtype=np.float64).reshape(-1, 3).astype(np.float32)
    sys_._write("POS", p.reshape(-1))


def _torch_functions():
    import torch

    class SoupEnergy(torch.autograd.Function):
        """U(x, theta) of a live system; backward gives dU/dx and dU/dtheta."""

        @staticmethod
        def forward(ctx, x, theta, system, basis):
            setPositions(system, x.detach().cpu().numpy())
            b = basisWith(basis, theta.detach().cpu().numpy(), system.headCharge)
            U = system.potential(b)["total"]
            res = energyAdjoint(system, b)
            ctx.save_for_backward(torch.as_tensor(res["grad"], device=x.device), torch.as_tensor(res["dU"], device=theta.device))
            return torch.tensor(U, dtype=torch.float64, device=x.device)

        @staticmethod
        def backward(ctx, gU):
            grad, dU = ctx.saved_tensors
            return gU * grad, gU * dU, None, None

    class SoupForce(torch.autograd.Function):
        """F(x, theta) = -dU/dx in FP64; backward is the force's vector-Jacobian product."""

        @staticmethod
        def forward(ctx, x, theta, system, basis):
            setPositions(system, x.detach().cpu().numpy())
            b = basisWith(basis, theta.detach().cpu().numpy(), system.headCharge)
            ctx.system, ctx.basis = system, b
            return torch.as_tensor(-energyAdjoint(system, b)["grad"], device=x.device)

        @staticmethod
        def backward(ctx, gF):
            res = energyAdjoint(ctx.system, ctx.basis, w=gF.detach().cpu().numpy())
            return (torch.as_tensor(-res["hw"], device=gF.device), torch.as_tensor(res["dFw"], device=gF.device), None, None)

    return SoupEnergy, SoupForce


def __getattr__(name):
    if name in ("SoupEnergy", "SoupForce"):
        e, f = _torch_functions()
        return e if name == "SoupEnergy" else f
    raise AttributeError(name)


# ---- trajectory segments -------------------------------------------------------------------------
def _topology(sys_, pos: np.ndarray, bonds: np.ndarray, links: np.ndarray, charges: np.ndarray) -> None:
    """Loads one recorded state into a system and rebuilds its lists (positions, bond slots,
    centre links, charges)."""
    sys_._write("POS", np.ascontiguousarray(pos, dtype=np.float32).reshape(-1))
    sys_._write("BOND_SLOTS", np.ascontiguousarray(bonds, dtype=np.uint32).reshape(-1))
    sys_._write("CENTER_LINK", np.ascontiguousarray(links, dtype=np.uint32).reshape(-1))
    sys_._write("CHARGE", np.ascontiguousarray(charges, dtype=np.float32).reshape(-1))
    sys_._call(sys_._lib.soup_forces, native.FORCES_REBUILD)


def _clone(sys_):
    """A second handle holding the same state and parameters (for the reverse sweep's launches)."""
    from .checkpoint import decodeCheckpointResume, encodeCheckpoint
    from .sim import createSoup

    R = sys_._R
    opts = dict(R.opts)
    opts["resume"] = decodeCheckpointResume(encodeCheckpoint(sys_, {}))
    c = createSoup(opts)
    c._call(c._lib.soup_set_params, C.byref(R.cfg))
    return c


def segmentGradient(sys_, steps: int, basis: Optional[native.SoupPotentialBasis] = None) -> Dict[str, Any]:
    """Advances the system by `steps` steps, exactly as step(steps) would, and returns the gradient of
    J = U(x_K; theta), the potential energy at the end, with respect to theta, to the start positions
    x_0 and to the start velocities v_0, by one reverse sweep over the recorded states.

    The step is the reference's velocity-Verlet Langevin step (step.wgsl:96-142): vh = v + dt/2 F(x),
    x' = x + dt vh, v' = (1 - gamma dt)(vh + dt/2 F(x')) + noise. The noise does not depend on the
    state, so the pathwise adjoint runs backward through the two half-kicks and the drift, with one
    adjoint launch (-H w and d(w . F)/dtheta) per force evaluation. Bond and charge Monte Carlo are
    not differentiated: each force is taken with the bond slots, centre links and charges it was
    computed with, as recorded; a constant-pH sweep inside the segment splits the force a step
    starts from from the one the previous step ended with, and both are honoured. The reverse
    launches run on a second handle, so the system continues exactly as step(steps) leaves it.
    The start must not follow a bond Monte Carlo step (the force buffer would hold the previous
    topology's force)."""
    R = sys_._R
    cfg = R.cfg
    n = sys_.N
    g0 = sys_.steps
    if g0 > 0 and (g0 - 1) % cfg.bond_interval == 0:
        raise ValueError("segmentGradient: the segment starts right after a bond Monte Carlo step; start one step later")
    if int((sys_.frozen() != 0).sum()) > 0:
        raise ValueError("segmentGradient: frozen particles are not differentiated")
    b = basis if basis is not None else sys_.potentialBasis()
    K = int(steps)
    x0 = sys_.particles().reshape(n, 4).copy()
    B0, L0, C0 = sys_.bondSlots().reshape(n, 3).copy(), sys_.centerLinks().copy(), sys_.charges().copy()
    P = np.empty((K, n, 4), np.float32)
    Bs = np.empty((K, n, 3), np.uint32)
    Ls = np.empty((K, n), np.uint32)
    Cs = np.empty((K, n), np.float32)
    sw = np.zeros(K, np.uint8)
    sys_._call(sys_._lib.soup_step_record, C.c_uint32(K), P.ctypes.data_as(C.c_void_p), Bs.ctypes.data_as(C.c_void_p),
               Ls.ctypes.data_as(C.c_void_p), Cs.ctypes.data_as(C.c_void_p), sw.ctypes.data_as(C.c_void_p))
    J = sys_.potential(b)["total"]
    end = energyAdjoint(sys_, b)
    dt, a = float(cfg.dt), 1.0 - float(cfg.gamma) * float(cfg.dt)
    clone = _clone(sys_)
    try:
        def start_of(s):  # topology at the start of step s
            return (B0, L0, C0) if s == 0 else (Bs[s - 1], Ls[s - 1], Cs[s - 1])

        def vjp(pos, topo, w):
            _topology(clone, pos, *topo)
            r = energyAdjoint(clone, b, w=w)
            return -r["hw"], r["dFw"]

        lx = end["grad"].copy()
        lv = np.zeros((n, 3))
        lth = end["dU"].copy()
        lvh_next = np.zeros((n, 3))
        for s in range(K - 1, -1, -1):
            wA = a * (dt / 2) * lv                       # step s's second half-kick
            wB = (dt / 2) * lvh_next if s < K - 1 else None  # step s + 1's first half-kick
            if wB is not None and sw[s]:
                gx, gt = vjp(P[s], (Bs[s], Ls[s], Cs[s]), wB)
                lx += gx
                lth += gt
                wB = None
            w = wA if wB is None else wA + wB
            gx, gt = vjp(P[s], start_of(s), w)
            lx += gx
            lth += gt
            lvh = a * lv + dt * lx
            lv = lvh
            lvh_next = lvh
        gx, gt = vjp(x0, (B0, L0, C0), (dt / 2) * lvh_next)  # step 0's first half-kick
        lx += gx
        lth += gt
    finally:
        clone.dispose()
    return {"J": J, "dJ_dtheta": lth, "dJ_dx0": lx, "dJ_dv0": lv, "steps": K, "sweeps": int(sw.sum())}

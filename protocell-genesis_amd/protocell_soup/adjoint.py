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

segmentGradient differentiates a trajectory observable over a segment of steps (soup_segment_adjoint:
device checkpoints every checkpointEvery steps, each stretch recomputed forward before its reverse
sweep). Its observables are the end energy and the window mean of a strain virial; the lateral
tension of a membrane normal to z is the one the area-per-lipid sensitivity goes through: at fixed
box the area per lipid a cannot move, but the tension gamma(a, theta) it holds can, and at the
tensionless state a barostat would find, da/dtheta = -(a / K_A) dgamma/dtheta with K_A = a dgamma/da
the area compressibility modulus (areaPerLipidSensitivity, areaCompressibility).
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


def energyAdjoint(sys_, basis: Optional[native.SoupPotentialBasis] = None, w: Optional[np.ndarray] = None,
                  strain: Optional[Sequence[float]] = None) -> Dict[str, Any]:
    """One adjoint launch at the system's configuration: grad = dU/dx (N, 3), dU = dU/dtheta, and with
    a direction w (N, 3): hw = H w and dFw = d(w . F)/dtheta. FP64 throughout.

    With a diagonal strain e (3 values) every pair term's direction also takes diag(e) r_ij: then
    We = W_e, the strain virial (sum over pair terms of (diag(e) r_ij) . f_ij), hw = H w + H_e and
    dFw = d(w . F)/dtheta + dW_e/dtheta, and dW_e/dx = -diag(e) grad - H_e (virialGradient)."""
    b = basis if basis is not None else sys_.potentialBasis()
    n = sys_.N
    grad = np.empty(3 * n, dtype=np.float64)
    dU = np.empty(14, dtype=np.float64)
    hw = dFw = None
    wp = hp = fp = None
    if w is not None or strain is not None:
        hw = np.empty(3 * n, dtype=np.float64)
        dFw = np.empty(14, dtype=np.float64)
        hp, fp = hw.ctypes.data_as(C.c_void_p), dFw.ctypes.data_as(C.c_void_p)
    if w is not None:
        wa = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        if wa.size != 3 * n:
            raise ValueError(f"energyAdjoint: w holds {wa.size // 3} particles, the system {n}")
        wp = wa.ctypes.data_as(C.c_void_p)
    out = {"grad": grad.reshape(n, 3), "dU": dU}
    if strain is None:
        sys_._call(sys_._lib.soup_energy_adjoint, C.byref(b), wp, grad.ctypes.data_as(C.c_void_p), hp,
                   dU.ctypes.data_as(C.c_void_p), fp)
    else:
        e = (C.c_double * 3)(*[float(x) for x in strain])
        We = C.c_double(0.0)
        sys_._call(sys_._lib.soup_energy_adjoint_strain, C.byref(b), e, wp, grad.ctypes.data_as(C.c_void_p), hp,
                   dU.ctypes.data_as(C.c_void_p), fp, C.byref(We))
        out["We"] = We.value
    if hw is not None:
        out["hw"] = hw.reshape(n, 3)
        out["dFw"] = dFw
    return out


def virialGradient(sys_, strain: Sequence[float], basis: Optional[native.SoupPotentialBasis] = None) -> Dict[str, Any]:
    """The strain virial W_e of the current configuration and its gradients: dx = dW_e/dx (N, 3) and
    dtheta = dW_e/dtheta."""
    r = energyAdjoint(sys_, basis, strain=strain)
    e = np.asarray(strain, dtype=np.float64)
    return {"W": r["We"], "dx": -r["grad"] * e - r["hw"], "dtheta": r["dFw"]}


def tensionStrain(sys_) -> tuple:
    """The strain and scale whose W_e is the configurational lateral tension of a membrane normal to
    z, (W_zz - (W_xx + W_yy) / 2) / (Lx Ly)."""
    L = sys_.box
    return (-0.5, -0.5, 1.0), 1.0 / (float(L[0]) * float(L[1]))


def areaCompressibility(areas: Sequence[float], kT: float) -> float:
    """K_A = kT <A> / var(A) from a trace of the projected area A sampled at zero tension (the area
    move's accepted boxes)."""
    a = np.asarray(areas, dtype=np.float64)
    return float(kT * a.mean() / a.var())


def areaPerLipidSensitivity(dtension_dtheta: np.ndarray, area_per_lipid: float, KA: float) -> np.ndarray:
    """da/dtheta = -(a / K_A) dgamma/dtheta: how far the tensionless area per lipid a moves per unit
    of each parameter, from the fixed-area tension sensitivity (implicit function theorem on
    gamma(a*(theta), theta) = 0)."""
    return -(float(area_per_lipid) / float(KA)) * np.asarray(dtension_dtheta, dtype=np.float64)


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
def segmentGradient(sys_, steps: int, basis: Optional[native.SoupPotentialBasis] = None, checkpointEvery: int = 200,
                    objective: str = "energy", strain: Optional[Sequence[float]] = None, scale: float = 1.0) -> Dict[str, Any]:
    """Advances the system by `steps` steps, exactly as step(steps) would, and returns the gradient of
    a trajectory observable J with respect to theta, to the start positions x_0 and to the start
    velocities v_0, from one reverse sweep (soup_segment_adjoint):
    - objective "energy": J = U(x_K; theta), the potential energy at the end;
    - "tension": J = the mean over the states after each step of the configurational lateral tension
      (tensionStrain), the observable of the area-per-lipid sensitivity;
    - "virial": J = scale / K * sum over those states of W_e for the given strain.

    The step is the reference's velocity-Verlet Langevin step (step.wgsl:96-142): vh = v + dt/2 F(x),
    x' = x + dt vh, v' = (1 - gamma dt)(vh + dt/2 F(x')) + noise. The noise does not depend on the
    state, so the pathwise adjoint runs backward through the two half-kicks and the drift with one
    force vector-Jacobian product per force evaluation, each with the bond slots, centre links and
    charges that force was computed with; Monte Carlo moves are not differentiated. The library
    keeps a checkpoint of the full state on the device every `checkpointEvery` steps and recomputes
    each stretch from it on a replay system before sweeping it backwards, so the device holds one
    stretch of states, not the whole segment. The start step must be a multiple of rebuildEvery and
    must not follow a bond Monte Carlo step; frozen particles are not differentiated."""
    b = basis if basis is not None else sys_.potentialBasis()
    spec = native.SoupSegmentSpec()
    spec.steps = int(steps)
    spec.checkpoint_every = int(checkpointEvery)
    if objective == "energy":
        spec.objective = native.OBJ_END_ENERGY
    elif objective in ("tension", "virial"):
        spec.objective = native.OBJ_MEAN_STRAIN_VIRIAL
        if objective == "tension":
            strain, scale = tensionStrain(sys_)
        elif strain is None:
            raise ValueError("segmentGradient: the virial objective needs a strain")
        for a in range(3):
            spec.strain[a] = float(strain[a])
        spec.scale = float(scale)
    else:
        raise ValueError(f"segmentGradient: objective {objective!r} is 'energy', 'tension' or 'virial'")
    n = sys_.N
    J = C.c_double(0.0)
    dth = np.empty(14, dtype=np.float64)
    dx0 = np.empty(3 * n, dtype=np.float64)
    dv0 = np.empty(3 * n, dtype=np.float64)
    sys_._call(sys_._lib.soup_segment_adjoint, C.byref(b), C.byref(spec), C.byref(J), dth.ctypes.data_as(C.c_void_p),
               dx0.ctypes.data_as(C.c_void_p), dv0.ctypes.data_as(C.c_void_p))
    return {"J": J.value, "dJ_dtheta": dth, "dJ_dx0": dx0.reshape(n, 3), "dJ_dv0": dv0.reshape(n, 3), "steps": int(steps),
            "objective": objective}

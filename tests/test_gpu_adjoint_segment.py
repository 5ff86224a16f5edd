"""The strain-virial adjoint and the trajectory-segment adjoint with device checkpoints on the GPU.

- soup_energy_adjoint_strain against the oracle's FP64 adjoint (or_energy_adjoint, pinned in
  test_oracle_adjoint_cpu.py by differences of the oracle's energy and by the affine-strain identity
  W_e = -dU/deps).
- soup_segment_adjoint: the stretches recomputed from device checkpoints reproduce the forward pass
  bit for bit, so the result does not depend on checkpoint_every, and the system ends exactly where
  step(K) leaves it; the gradient equals the oracle's FP64 reverse sweep over the same recorded
  trajectory (segment_gradient_f64).
- The area-per-lipid sensitivity on the 2,087-amphiphile bilayer patch (39,999 beads, charged heads
  at 0.01 M) goes through the window-mean lateral tension: the 200-step window's gradient against the
  oracle's FP64 reverse sweep, a 50-step window's against central differences of reruns (step 1e-2;
  the chaotic f32 dynamics carry the reruns out of the linear regime before 200 steps), and the
  2,000-step protocol with a checkpoint every 200 steps, as ten windows and as one pathwise segment.
"""
import ctypes as C
import math
import time

import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup import native
from protocell_soup.adjoint import (CLASS_PAIRS, PARAM_NAMES, areaPerLipidSensitivity, basisWith, energyAdjoint,
                                    paramsOf, segmentGradient, tensionStrain, virialGradient)
from soup_fixtures import amphiphile_opts, bilayer_patch_opts, oracle_from_system

pytestmark = pytest.mark.gpu

CHARGED = {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1}


def _rel_max(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _small(steps=70):
    s = ps.createSoup(amphiphile_opts(16.0, 80, 4, {"W": 1800, "C": 150}, seed=6, electrostatics=CHARGED))
    s.relaxColdStart()
    s.step(steps)  # a multiple of rebuildEvery, not right after a bond Monte Carlo step
    return s


def test_strain_adjoint_matches_oracle_fp64():
    s = _small(60)
    try:
        b = s.potentialBasis()
        e = (0.3, -0.7, 1.1)
        w = np.random.default_rng(4).normal(size=(s.N, 3))
        r = energyAdjoint(s, b, w=w, strain=e)
        o = oracle_from_system(s, s._R)
        o.rebuild()
        ro = o.energy_adjoint(b, _box(s), s.headCharge, w=w, strain=e)
        assert r["We"] == pytest.approx(ro["We"], rel=1e-11)
        assert _rel_max(r["grad"], ro["grad"]) < 1e-11
        assert _rel_max(r["hw"], ro["hw"]) < 1e-11
        assert _rel_max(r["dFw"], ro["dFw"]) < 1e-11 and _rel_max(r["dU"], ro["dU"]) < 1e-11
        # the strain alone: W_e and its gradients
        v = virialGradient(s, e, b)
        rv = o.energy_adjoint(b, _box(s), s.headCharge, strain=e)
        assert v["W"] == pytest.approx(rv["We"], rel=1e-11)
        assert _rel_max(v["dx"], -rv["grad"] * np.asarray(e) - rv["hw"]) < 1e-11
        assert _rel_max(v["dtheta"], rv["dFw"]) < 1e-11
        # and without a strain the call is soup_energy_adjoint
        r0 = energyAdjoint(s, b, w=w)
        r1 = energyAdjoint(s, b, w=w, strain=(0.0, 0.0, 0.0))
        assert np.array_equal(r0["hw"], r1["hw"]) and np.array_equal(r0["dFw"], r1["dFw"]) and r1["We"] == 0.0
    finally:
        s.dispose()


def _state(s):
    return (s.particles().copy(), s.velocities().copy(), s.bondSlots().copy(), s.charges().copy(), s.rngState()["thermo"].copy())


def test_segment_checkpoints_recompute_the_forward_pass():
    K = 60
    runs = []
    for C_ in (10, 20, 60):
        s = _small()
        try:
            b = s.potentialBasis()
            res = segmentGradient(s, K, b, checkpointEvery=C_, objective="tension")
            runs.append((res, _state(s), s.steps))
        finally:
            s.dispose()
    t = _small()
    try:
        t.step(K)
        plain = _state(t)
    finally:
        t.dispose()
    for res, st, steps in runs:
        assert steps == 70 + K
        for a, c in zip(st, plain):
            assert np.array_equal(a, c)
        r0 = runs[0][0]
        assert res["J"] == r0["J"]
        assert np.array_equal(res["dJ_dtheta"], r0["dJ_dtheta"])
        assert np.array_equal(res["dJ_dx0"], r0["dJ_dx0"]) and np.array_equal(res["dJ_dv0"], r0["dJ_dv0"])
    assert np.isfinite(runs[0][0]["dJ_dtheta"]).all() and np.abs(runs[0][0]["dJ_dtheta"]).max() > 0


def _record(s, K):
    """K steps of s with the state after every step (soup_step_record) and the state before."""
    n = s.N
    start = (s.particles().copy(), s.bondSlots().copy(), s.centerLinks().copy(), s.charges().copy())
    P = np.empty((K, 4 * n), np.float32)
    B = np.empty((K, 3 * n), np.uint32)
    Lk = np.empty((K, n), np.uint32)
    Q = np.empty((K, n), np.float32)
    sw = np.zeros(K, np.uint8)
    s._call(s._lib.soup_step_record, C.c_uint32(K), P.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p),
            Lk.ctypes.data_as(C.c_void_p), Q.ctypes.data_as(C.c_void_p), sw.ctypes.data_as(C.c_void_p))
    return start, [(P[k], B[k], Lk[k], Q[k]) for k in range(K)], sw


def _box(s):
    """The box edges as the device holds them (f32): soup_potential and the adjoint take the minimum
    image over these, so the oracle does too (a box edge that f32 cannot hold, 36.84, would otherwise
    move each boundary-crossing pair by its rounding)."""
    return np.asarray(s._R.cfg.box, dtype=np.float64)


def _oracle_sweep(s, b, start, recs, sw, g0, objective, strain=None, scale=1.0):
    R = s._R
    o = oracle_from_system(s, R)
    cfg = R.cfg
    return orc.segment_gradient_f64(o, b, _box(s), s.headCharge, float(cfg.dt), float(cfg.gamma), start, recs, sw,
                                    int(cfg.bond_interval), g0, objective=objective, strain=strain, scale=scale)


@pytest.mark.parametrize("objective", ["energy", "tension"])
def test_segment_gradient_matches_oracle_sweep(objective):
    K = 40
    s = _small()
    try:
        b = s.potentialBasis()
        res = segmentGradient(s, K, b, checkpointEvery=20, objective=objective)
    finally:
        s.dispose()
    t = _small()
    try:
        strain, scale = tensionStrain(t)
        start, recs, sw = _record(t, K)
        ref = _oracle_sweep(t, b, start, recs, sw, 70, "energy" if objective == "energy" else "virial", strain, scale)
    finally:
        t.dispose()
    assert res["J"] == pytest.approx(ref["J"], rel=1e-10)
    assert _rel_max(res["dJ_dtheta"], ref["dJ_dtheta"]) < 1e-8
    assert _rel_max(res["dJ_dx0"], ref["dJ_dx0"]) < 1e-8 and _rel_max(res["dJ_dv0"], ref["dJ_dv0"]) < 1e-8


# ---- the bilayer patch: area per lipid through the lateral tension ---------------------------------
WINDOW = 200


def _patch():
    s = ps.createSoup(bilayer_patch_opts())
    s.relaxColdStart({"iterations": 200})
    s.step(400)  # a multiple of rebuildEvery, not right after a bond Monte Carlo step
    return s


def _set_theta(s, b, k, value):
    """Moves parameter k of the device's force law to `value` (as f32) and returns the basis at the
    value the device holds."""
    cfg = native.SoupConfig()
    C.memmove(C.byref(cfg), C.byref(s._R.cfg), C.sizeof(native.SoupConfig))
    if k < 10:
        a, c = CLASS_PAIRS[k]
        cfg.attr[a][c] = value
        cfg.attr[c][a] = cfg.attr[a][c]
        held = float(cfg.attr[a][c])
    elif k == 11:
        cfg.epsilon = value
        held = float(cfg.epsilon)
    elif k == 13:
        cfg.es_kappa = value
        held = float(cfg.es_kappa)
        rc = float(cfg.es_rc)
        cfg.es_Frc = float(cfg.es_A) * math.exp(-held * rc) / rc * (1.0 / rc + held)
    else:
        raise ValueError(k)
    s._call(s._lib.soup_set_params, C.byref(cfg))
    th = paramsOf(b, s.headCharge)
    th[k] = held
    return basisWith(b, th, s.headCharge), held


def test_bilayer_window_gradient_matches_oracle_fp64():
    """The 200-step window's gradient (one stretch, its checkpoint at the window start) against the
    oracle's FP64 reverse sweep over the same recorded trajectory."""
    t0 = time.time()
    s = _patch()
    try:
        assert s.N == 39999 and s.steps == 400
        b = s.potentialBasis()
        res = segmentGradient(s, WINDOW, b, checkpointEvery=200, objective="tension")
        print(f"PATCH window {WINDOW}: tension {res['J']:.6f}, adjoint in {time.time() - t0:.1f} s", flush=True)
    finally:
        s.dispose()
    assert np.isfinite(res["dJ_dtheta"]).all()
    t = _patch()
    try:
        strain, scale = tensionStrain(t)
        start, recs, sw = _record(t, WINDOW)
        t1 = time.time()
        ref = _oracle_sweep(t, b, start, recs, sw, 400, "virial", strain, scale)
        print(f"PATCH oracle sweep {time.time() - t1:.1f} s", flush=True)
    finally:
        t.dispose()
    assert res["J"] == pytest.approx(ref["J"], rel=1e-10)
    assert _rel_max(res["dJ_dtheta"], ref["dJ_dtheta"]) < 1e-8
    assert _rel_max(res["dJ_dx0"], ref["dJ_dx0"]) < 1e-8 and _rel_max(res["dJ_dv0"], ref["dJ_dv0"]) < 1e-8


def _fd(b, k, W, rel):
    th0 = paramsOf(b, -1.0)
    h = rel * max(1.0, abs(th0[k]))
    Js, held = [], []
    for sgn in (1.0, -1.0):
        r = _patch()
        try:
            bk, v = _set_theta(r, b, k, th0[k] + sgn * h)
            held.append(v)
            Js.append(segmentGradient(r, W, bk, checkpointEvery=200, objective="tension")["J"])
        finally:
            r.dispose()
    return (Js[0] - Js[1]) / (held[0] - held[1])


def test_bilayer_window_gradient_against_differences():
    """Central differences of reruns (parameter moved by +-1e-2 relative) against the pathwise adjoint
    on a 50-step window, for the class depths with a well (classes: 0 tails, 1 heads, 2 water;
    tail-head and head-head have none, and a zero depth is a kink: soupPotential skips pairs of depth
    <= 0), epsilon and kappa. The reruns integrate in f32 and the dynamics are chaotic: a perturbation
    of the parameters grows along the trajectory until the reruns leave the linear regime, which on
    this patch happens between 100 and 200 steps; at 200 steps the differences no longer estimate the
    pathwise derivative (printed, not asserted), which the oracle's FP64 sweep checks instead."""
    W = 50
    s = _patch()
    try:
        b = s.potentialBasis()
        res = segmentGradient(s, W, b, checkpointEvery=200, objective="tension")
    finally:
        s.dispose()
    g = res["dJ_dtheta"]
    names = ("attr00", "attr12", "attr22", "epsilon", "kappa")
    # the bar: 1e-4 of the component, or of the largest tested component for a small one (the f32
    # reruns carry a noise floor of their own, about 1e-3 absolute here)
    gscale = max(abs(g[PARAM_NAMES.index(n)]) for n in names)
    for name in names:
        k = PARAM_NAMES.index(name)
        fd = _fd(b, k, W, 1e-2)
        print(f"PATCH {W} steps {name}: adjoint {g[k]:.8e} differences {fd:.8e} rel {abs(fd - g[k]) / abs(g[k]):.2e}",
              flush=True)
        assert abs(fd - g[k]) <= 1e-4 * max(abs(g[k]), gscale), name
    s = _patch()
    try:
        g200 = segmentGradient(s, WINDOW, b, checkpointEvery=200, objective="tension")["dJ_dtheta"]
    finally:
        s.dispose()
    k = PARAM_NAMES.index("attr22")
    print(f"PATCH {WINDOW} steps attr22: adjoint {g200[k]:.6e} differences {_fd(b, k, WINDOW, 1e-2):.6e} (chaotic regime)",
          flush=True)


def test_bilayer_tension_sensitivity_over_2000_steps():
    """The c5 protocol: 2,000 f32 steps of the patch with a checkpoint every 200. The sensitivity of
    the long-run tension is the mean of the ten 200-step windows' gradients, each window's start
    state held fixed (the windowed, ensemble adjoint of a chaotic system); the single 2,000-step
    pathwise gradient through all ten recomputed stretches is computed as well, and the system ends
    exactly where 2,000 plain steps leave it."""
    s = _patch()
    try:
        b = s.potentialBasis()
        t0 = time.time()
        wins = [segmentGradient(s, WINDOW, b, checkpointEvery=200, objective="tension") for _ in range(10)]
        t_win = time.time() - t0
        end_w = _state(s)
    finally:
        s.dispose()
    s = _patch()
    try:
        t0 = time.time()
        full = segmentGradient(s, 2000, b, checkpointEvery=200, objective="tension")
        t_full = time.time() - t0
        end_f = _state(s)
    finally:
        s.dispose()
    # step(n) runs the constant-pH sweep at its chunk ends once one is due, so ten step(200) calls and
    # one step(2000) sweep at different steps; each adjoint run ends where its plain twin does
    for calls, end in (((200,) * 10, end_w), ((2000,), end_f)):
        t = _patch()
        try:
            for n in calls:
                t.step(n)
            plain = _state(t)
        finally:
            t.dispose()
        for a, c in zip(end, plain):
            assert np.array_equal(a, c)
    G = np.array([w["dJ_dtheta"] for w in wins])
    J = np.array([w["J"] for w in wins])
    assert np.isfinite(G).all() and np.isfinite(full["dJ_dtheta"]).all()
    # the windows' mean tension against the single segment's (their trajectories part at the sweep)
    print(f"PATCH tension: windows {J.mean():.6f}, one segment {full['J']:.6f}", flush=True)
    mean, sem = G.mean(axis=0), G.std(axis=0, ddof=1) / math.sqrt(len(G))
    a = 36.84 ** 2 / (2087 / 2)
    for name in ("attr00", "attr12", "attr22", "epsilon", "headCharge", "kappa"):
        k = PARAM_NAMES.index(name)
        print(f"PATCH 2000 steps {name}: dtension/dtheta windows {mean[k]:+.4e} +- {sem[k]:.1e}, "
              f"pathwise {full['dJ_dtheta'][k]:+.4e}; d(area per lipid)/dtheta x K_A {areaPerLipidSensitivity(mean, a, 1.0)[k]:+.4e}",
              flush=True)
    print(f"PATCH 2000 steps: tension {full['J']:.5f}, ten windows {t_win:.1f} s, one pathwise segment {t_full:.1f} s", flush=True)

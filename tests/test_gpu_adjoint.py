"""The energy and force adjoint (soup_energy_adjoint) against finite differences of the GPU's own
FP64 energy, against the oracle's FP64 force and its CPU energy, and through the torch.autograd
front. The bar is 1e-4 relative; the checks below hold far tighter.

Finite differences in the positions use the displacement the device actually holds (the f32
rounding of x +- h d), so the comparison is U(x+) - U(x-) against grad . (x+ - x-), exact to third
order, rather than a difference quotient that f32 rounding would blur."""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.adjoint import PARAM_NAMES, SoupEnergy, SoupForce, basisWith, paramsOf, setPositions
from oracle import oracle as orc
from soup_fixtures import amphiphile_opts, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

CHARGED = {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1}


@pytest.fixture(scope="module")
def soup():
    s = ps.createSoup(amphiphile_opts(16.0, 80, 4, {"W": 1800, "C": 150, "M": 10}, seed=6, electrostatics=CHARGED))
    s.relaxColdStart()
    s.step(60)
    yield s
    s.dispose()


def _x(s):
    return s.particles().reshape(-1, 4)[:, :3].astype(np.float64)


def _U(s, b):
    return s.potential(b)["total"]


def _active_params(s, b):
    """Parameters the configuration's energy depends on smoothly: attraction pairs with a positive
    depth (a zero depth is a kink: soupPotential skips pairs of depth <= 0) and the rest."""
    th = paramsOf(b, s.headCharge)
    return [k for k in range(14) if k >= 10 or th[k] > 0.0]


def test_gradient_is_minus_the_force(soup):
    s = soup
    b = s.potentialBasis()
    res = s.energyAdjoint(b)
    o = oracle_from_system(s, s._R)
    f64 = o.force_f64(use_list=False)
    # the step's force law takes its parameters as f32 (SoupConfig) and soupPotential as doubles
    # (the basis); the steep WCA core turns those ~1e-8 differences into ~5e-7 here
    assert force_rel_err(-res["grad"], f64) < 1e-6
    # and the step's f32 force within its own 1e-5 bound
    f32 = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
    assert force_rel_err(f32, -res["grad"]) < 1e-5


def test_position_gradient_and_hessian_against_differences(soup):
    s = soup
    b = s.potentialBasis()
    x0 = _x(s)
    rng = np.random.default_rng(2)
    d = rng.normal(size=x0.shape)
    res = s.energyAdjoint(b, w=d)
    errs, meds, bad = [], [], []
    try:
        for h in (2e-4, 1e-4, 5e-5):
            xp = (x0 + h * d).astype(np.float32).astype(np.float64)
            xm = (x0 - h * d).astype(np.float32).astype(np.float64)
            setPositions(s, xp)
            up, gp = _U(s, b), s.energyAdjoint(b)["grad"]
            setPositions(s, xm)
            um, gm = _U(s, b), s.energyAdjoint(b)["grad"]
            dx = xp - xm
            lin = float((res["grad"] * dx).sum())
            # against the scale of the first-order change (its terms before they cancel)
            scale = float(np.abs(res["grad"] * dx).sum())
            errs.append(abs((up - um) - lin))
            assert errs[-1] <= 1e-6 * scale, (h, errs[-1], scale)
            # H (x+ - x-) from the adjoint's Hessian-vector product (linear in w), per particle. The
            # second derivatives jump where a pair crosses the WCA cut or an edge of the attraction
            # shell, so the few particles with such a pair between x- and x+ differ by that jump:
            # the rest must agree to the truncation of the difference
            setPositions(s, x0)
            hv = s.energyAdjoint(b, w=dx)["hw"]
            n_i = np.linalg.norm(hv, axis=1)
            err_i = np.linalg.norm((gp - gm) - hv, axis=1) / np.maximum(n_i, np.median(n_i))
            meds.append(float(np.median(err_i)))
            bad.append(float((err_i > 1e-4).mean()))
        # the remainders: the truncation (~h^2), the f32 rounding of x+- around x0 (the midpoint
        # is not x0 to ~5e-7, a floor independent of h), and the particles with a pair across a
        # kink, fewer at smaller h; a wrong second derivative would miss by O(1) everywhere
        print("energy residuals", errs, "hessian median errors", meds, "fraction above 1e-4", bad)
        assert max(meds) < 2e-5 and bad[2] < 0.02 and bad[2] < bad[0], (meds, bad)
    finally:
        setPositions(s, x0)


def test_parameter_gradients_against_gpu_and_cpu_differences(soup):
    s = soup
    b = s.potentialBasis()
    th0 = paramsOf(b, s.headCharge)
    rng = np.random.default_rng(5)
    w = rng.normal(size=(s.N, 3))
    res = s.energyAdjoint(b, w=w)
    o = oracle_from_system(s, s._R)
    pos, bs, cl, q = s.particles(), s.bondSlots(), s.centerLinks(), s.charges()
    active = _active_params(s, b)
    assert {10, 11, 12, 13} <= set(active) and len(active) >= 10
    for k in active:
        h = 1e-4 * max(1.0, abs(th0[k]))  # U is linear or quadratic in all but kappa; large h keeps FP64 cancellation small
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        bp, bm = basisWith(b, tp, s.headCharge), basisWith(b, tm, s.headCharge)
        # GPU energy
        fd = (_U(s, bp) - _U(s, bm)) / (2 * h)
        assert fd == pytest.approx(res["dU"][k], rel=1e-6, abs=1e-7), PARAM_NAMES[k]
        # the oracle's FP64 CPU energy
        cp = orc.soup_potential(bp, pos, bs, cl, q, s.box)[4]
        cm = orc.soup_potential(bm, pos, bs, cl, q, s.box)[4]
        assert (cp - cm) / (2 * h) == pytest.approx(res["dU"][k], rel=1e-6, abs=1e-7), PARAM_NAMES[k]
        # d(w . F)/dtheta against the difference of w . F = -w . grad
        wfp = -float((s.energyAdjoint(bp)["grad"] * w).sum())
        wfm = -float((s.energyAdjoint(bm)["grad"] * w).sum())
        assert (wfp - wfm) / (2 * h) == pytest.approx(res["dFw"][k], rel=1e-5, abs=1e-6), PARAM_NAMES[k]
    del o


def test_torch_energy_and_force_functions(soup):
    torch = pytest.importorskip("torch")
    s = soup
    b = s.potentialBasis()
    x0 = _x(s)
    th0 = paramsOf(b, s.headCharge)
    try:
        x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
        th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
        U = SoupEnergy.apply(x, th, s, b)
        U.backward()
        ref = s.energyAdjoint(b)
        assert U.item() == pytest.approx(_U(s, b), rel=1e-14)
        assert np.array_equal(x.grad.numpy(), ref["grad"]) and np.array_equal(th.grad.numpy(), ref["dU"])
        # a loss on the forces: L = sum(c * F); dL/dx = -H c, dL/dtheta = d(c . F)/dtheta
        c = torch.tensor(np.random.default_rng(9).normal(size=x0.shape), dtype=torch.float64)
        x2 = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
        th2 = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
        L = (SoupForce.apply(x2, th2, s, b) * c).sum()
        L.backward()
        vjp = s.energyAdjoint(b, w=c.numpy())
        assert np.array_equal(x2.grad.numpy(), -vjp["hw"]) and np.array_equal(th2.grad.numpy(), vjp["dFw"])
        # and against the difference of L in a parameter (smooth, unlike a position move across the
        # WCA cut): the DH kappa
        k = PARAM_NAMES.index("kappa")
        h = 1e-6
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        Lp = float((SoupForce.apply(torch.tensor(x0), torch.tensor(tp), s, b) * c).sum())
        Lm = float((SoupForce.apply(torch.tensor(x0), torch.tensor(tm), s, b) * c).sum())
        assert (Lp - Lm) / (2 * h) == pytest.approx(float(th2.grad[k]), rel=1e-5)
    finally:
        setPositions(s, x0)


def _segment_soup():
    s = ps.createSoup(amphiphile_opts(16.0, 80, 4, {"W": 1800}, seed=6, electrostatics=CHARGED))
    s.relaxColdStart()
    s.step(70)  # a multiple of rebuildEvery, not right after a bond Monte Carlo step (every 20)
    return s


def test_segment_gradient_against_perturbed_reruns():
    import ctypes as C

    from protocell_soup import native
    from protocell_soup.adjoint import segmentGradient

    K = 20  # one bond Monte Carlo pass inside
    s = _segment_soup()
    try:
        b = s.potentialBasis()
        res = segmentGradient(s, K, b)
        assert s.steps == 70 + K
        bonds_end = s.bondSlots()
        pos_end = s.particles()
    finally:
        s.dispose()
    # the recording run is the plain trajectory: the same end state as step(K)
    t = _segment_soup()
    try:
        t.step(K)
        assert np.array_equal(t.particles(), pos_end)
        assert t.potential(b)["total"] == res["J"]
    finally:
        t.dispose()
    # d J / d theta against reruns of the device dynamics with the parameter moved by +-h (a power
    # of two, so the f32 value on the device moves by exactly h). The reruns integrate in f32, whose
    # rounding the dynamics amplify (measured at 40 steps: 5e-4 on the tail pair, 3e-5 on the water
    # pair); at 20 steps and h = 2^-8 that noise and the truncation stay near 1e-5
    h = 2.0 ** -8
    for k in (PARAM_NAMES.index("attr22"), PARAM_NAMES.index("attr00"), PARAM_NAMES.index("epsilon")):
        J = []
        for sign in (1.0, -1.0):
            r = _segment_soup()
            try:
                cfg = native.SoupConfig()
                C.memmove(C.byref(cfg), C.byref(r._R.cfg), C.sizeof(native.SoupConfig))
                th = paramsOf(b, r.headCharge)
                th[k] += sign * h
                if k == PARAM_NAMES.index("epsilon"):
                    cfg.epsilon = cfg.epsilon + sign * h
                else:
                    a_, c_ = [(x, y) for x in range(4) for y in range(x, 4)][k]
                    cfg.attr[a_][c_] = cfg.attr[a_][c_] + sign * h
                    cfg.attr[c_][a_] = cfg.attr[a_][c_]
                r._call(r._lib.soup_set_params, C.byref(cfg))
                r.step(K)
                assert np.array_equal(r.bondSlots(), bonds_end)  # no Monte Carlo event moved
                J.append(r.potential(basisWith(b, th, r.headCharge))["total"])
            finally:
                r.dispose()
        fd = (J[0] - J[1]) / (2 * h)
        print(PARAM_NAMES[k], "adjoint", res["dJ_dtheta"][k], "differences", fd)
        assert fd == pytest.approx(res["dJ_dtheta"][k], rel=1e-4), PARAM_NAMES[k]

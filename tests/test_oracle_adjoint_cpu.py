"""The oracle's FP64 adjoint (or_energy_adjoint) pinned by its own FP64 energy (or_soup_potential),
which restates soup/src/soup-potential.ts:145-301 and is pinned by closed forms in
test_oracle_cpu.py:

- the gradient equals central differences of the energy in the positions;
- H w equals central differences of the gradient along w;
- dU/dtheta equals central differences of the energy in each parameter, and d(w . F)/dtheta those
  of w . F;
- the strain virial W_e equals -dU/deps for the affine deformation x -> (1 + eps diag(e)) x of the
  positions and the box: a closed identity of pair potentials under the minimum image, independent
  of the pair-term derivatives the adjoint sums.

The positions of the strained configurations are kept in FP64 (or_soup_potential_d), since f32
would round the small affine moves. The system is a charged amphiphile soup with bonds, bends and
DH pairs, so every term type takes part."""
import numpy as np
import pytest

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup.adjoint import PARAM_NAMES, basisWith, paramsOf
from protocell_soup.potential import makePotentialBasis
from soup_fixtures import amphiphile_opts

CHARGED = {"enabled": True, "pH": 5.0, "ionicStrengthMolar": 0.1}


@pytest.fixture(scope="module")
def system():
    opts = amphiphile_opts(16.0, 60, 4, {"W": 1800, "C": 120}, seed=4, electrostatics=CHARGED)
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    assert o.relax(200, 0.1) == 0
    o.rebuild()
    b = makePotentialBasis(R.soup, R.p, R.es, centerLink=True, charges=True)
    q = float(R.cfg.ph_q)
    assert (o.charge != 0).sum() > 10 and q != 0.0
    return R, o, b, q


def _U(o, b, box, pos=None):
    p = o.pos if pos is None else pos
    return orc.soup_potential(b, p, o.bond_slots, o.center_link, o.charge, box)[4]


def test_gradient_and_hessian_against_differences(system):
    R, o, b, q = system
    box = np.asarray(R.cfg.box, dtype=np.float64)
    rng = np.random.default_rng(3)
    w = rng.normal(size=(o.n, 3))
    r = o.energy_adjoint(b, box, q, w=w)
    x0 = o.pos.copy()
    picks = rng.choice(o.n, 6, replace=False)
    h = 2.0 ** -12  # exact in f32 for coordinates below 16
    try:
        for i in picks:
            for a in range(3):
                vals = []
                for k in (-2, -1, 1, 2):
                    o.pos[:] = x0
                    o.pos[4 * i + a] = x0[4 * i + a] + np.float32(k * h)
                    vals.append(_U(o, b, box))
                fd = (8 * (vals[2] - vals[1]) - (vals[3] - vals[0])) / (12 * h)
                assert fd == pytest.approx(r["grad"][i, a], rel=1e-5, abs=1e-5 * max(1.0, np.abs(r["grad"][i]).max()))
        # H w against the gradient along w (the lists stay valid for these tiny moves)
        hh = 1e-4
        xp = (x0.reshape(-1, 4)[:, :3].astype(np.float64) + hh * w).astype(np.float32)
        xm = (x0.reshape(-1, 4)[:, :3].astype(np.float64) - hh * w).astype(np.float32)
        gs = []
        for xx in (xp, xm):
            o.pos[:] = x0
            o.pos.reshape(-1, 4)[:, :3] = xx
            gs.append(o.energy_adjoint(b, box, q)["grad"])
        dx = xp.astype(np.float64) - xm.astype(np.float64)
        o.pos[:] = x0
        hv = o.energy_adjoint(b, box, q, w=dx)["hw"]
        n_i = np.linalg.norm(hv, axis=1)
        err = np.linalg.norm((gs[0] - gs[1]) - hv, axis=1) / np.maximum(n_i, np.median(n_i))
        # a pair crossing a WCA cut or a shell edge between x- and x+ jumps in its second derivative
        assert np.median(err) < 1e-5 and (err > 1e-3).mean() < 0.02, (np.median(err), (err > 1e-3).mean())
    finally:
        o.pos[:] = x0


def test_parameter_derivatives_against_differences(system):
    R, o, b, q = system
    box = np.asarray(R.cfg.box, dtype=np.float64)
    w = np.random.default_rng(8).normal(size=(o.n, 3))
    r = o.energy_adjoint(b, box, q, w=w)
    th0 = paramsOf(b, q)
    active = [k for k in range(14) if k >= 10 or th0[k] > 0.0]
    assert len(active) >= 10
    for k in active:
        h = 1e-4 * max(1.0, abs(th0[k]))
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        bp, bm = basisWith(b, tp, q), basisWith(b, tm, q)
        fd = (_U(o, bp, box) - _U(o, bm, box)) / (2 * h)
        assert fd == pytest.approx(r["dU"][k], rel=1e-6, abs=1e-7), PARAM_NAMES[k]
        wfp = -float((o.energy_adjoint(bp, box, q)["grad"] * w).sum())
        wfm = -float((o.energy_adjoint(bm, box, q)["grad"] * w).sum())
        assert (wfp - wfm) / (2 * h) == pytest.approx(r["dFw"][k], rel=1e-5, abs=1e-6), PARAM_NAMES[k]


@pytest.mark.parametrize("strain", [(1.0, 0.0, 0.0), (-0.5, -0.5, 1.0), (0.3, -0.7, 1.1)])
def test_strain_virial_is_minus_the_affine_energy_derivative(system, strain):
    R, o, b, q = system
    box = np.asarray(R.cfg.box, dtype=np.float64)
    e = np.asarray(strain, dtype=np.float64)
    r = o.energy_adjoint(b, box, q, strain=e)
    x0 = o.pos.reshape(-1, 4).astype(np.float64)
    vals = []
    eps = 1e-5
    for k in (-2, -1, 1, 2):
        s = 1.0 + k * eps * e
        xs = x0.copy()
        xs[:, :3] *= s
        vals.append(orc.soup_potential_d(b, xs, o.bond_slots, o.center_link, o.charge, box * s)[4])
    dUde = (8 * (vals[2] - vals[1]) - (vals[3] - vals[0])) / (12 * eps)
    assert r["We"] == pytest.approx(-dUde, rel=1e-7), (strain, r["We"], -dUde)
    # the strain's position gradient -diag(e) g - H_e against differences of W_e in single coordinates
    # (h exact in f32; a stencil rarely straddles a WCA cut or a shell edge, where the second
    # derivative jumps)
    dWdx = -r["grad"] * e - r["hw"]
    xs0 = o.pos.copy()
    h = 2.0 ** -12
    try:
        for i in np.random.default_rng(11).choice(o.n, 6, replace=False):
            for a in range(3):
                vals = []
                for k in (-2, -1, 1, 2):
                    o.pos[:] = xs0
                    o.pos[4 * i + a] = xs0[4 * i + a] + np.float32(k * h)
                    vals.append(o.energy_adjoint(b, box, q, strain=e)["We"])
                fd = (8 * (vals[2] - vals[1]) - (vals[3] - vals[0])) / (12 * h)
                assert fd == pytest.approx(dWdx[i, a], rel=1e-5, abs=1e-5 * max(1.0, np.abs(dWdx[i]).max())), (i, a)
    finally:
        o.pos[:] = xs0

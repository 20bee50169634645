"""Pins of the oracle's positivity-preserving change of variables (-cond_sys,
problem_def.hpp:36-70; fun_implicit petsc_solver_impl.hpp:626-655; complex-step
MatMult :297-326, :386-401).

  * prim_to_cons and cons_to_prim are inverse maps (closed form);
  * the complex-step product through the map equals sigma dC v + J_R(C) dC v,
    with dC v from the closed-form derivative of the map and J_R from the
    oracle's plain complex step at C(W), and it matches central differences
    of the implicit function F(W) = sigma (C(W) - C(W_n)) + R(C(W));
  * F = U_t + R(C(W)); impose_ic stores cons_to_prim of the conservative IC;
    the L2 error converts back first.
"""
import numpy as np
import pytest

import cases as C
from pyoracle import Oracle


def dC(w, v, dims):
    """Closed-form (dC/dW) v per node."""
    if dims == 2:
        W, V = w.reshape(-1, 4), v.reshape(-1, 4)
        rho, u, vv, e = np.exp(W[:, 0]), W[:, 1], W[:, 2], np.exp(W[:, 3])
        E = e + 0.5 * (u * u + vv * vv)
        out = np.stack([rho * V[:, 0],
                        rho * (u * V[:, 0] + V[:, 1]),
                        rho * (vv * V[:, 0] + V[:, 2]),
                        rho * (E * V[:, 0] + u * V[:, 1] + vv * V[:, 2] + e * V[:, 3])], axis=1)
        return out.reshape(-1)
    raise NotImplementedError


CASES = {
    "vortex_p3": C.vortex_cfg(4, 3, 3),
    "bump_p3": C.bump_cfg(4, 2, 3),
    "explosion_vdw_hllc_p2": C.explosion_cfg(5, 4, 2, gas=1, riemann=1),
    "sod1d_p3": C.sod_cfg(12, 3),
}


def prim_state(o, cfg, seed=3):
    return o.cons_to_prim(C.random_state(cfg, seed))


@pytest.mark.parametrize("name", sorted(CASES))
def test_maps_are_inverse(name):
    cfg = dict(CASES[name], cond_sys=1)
    o = Oracle(cfg)
    c = C.random_state(cfg)
    w = o.cons_to_prim(c)
    np.testing.assert_allclose(o.prim_to_cons(w), c, rtol=1e-14, atol=1e-14)
    if cfg["dims"] == 2:
        W = w.reshape(-1, 4)
        Cc = c.reshape(-1, 4)
        np.testing.assert_allclose(np.exp(W[:, 0]), Cc[:, 0], rtol=1e-14)
        np.testing.assert_allclose(W[:, 1], Cc[:, 1] / Cc[:, 0], rtol=1e-14)


@pytest.mark.parametrize("name", ["vortex_p3", "bump_p3", "explosion_vdw_hllc_p2"])
def test_cstep_through_the_map(name):
    cfg = dict(CASES[name], cond_sys=1, tangent_bc_mode=1)
    o = Oracle(cfg)
    o0 = Oracle(dict(cfg, cond_sys=0))
    w = prim_state(o, cfg)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(w.size)
    sigma = 3.0
    got = o.jvp_cstep(w, v, sigma=sigma)
    c, cv = o.prim_to_cons(w), dC(w, v, 2)
    want = sigma * cv + o0.jvp_cstep(c, cv, sigma=0.0)
    assert C.rel_l2(got, want) <= 1e-9
    if name == "bump_p3":
        return  # complex step sees the lambda-bound wall frozen: no FD comparison
    # central differences of F(W) = sigma (C(W) - C(W_n)) + R(C(W))
    wn = w * 0.99
    cn = o.prim_to_cons(wn)

    def F(x):
        return o.ifunction(x, sigma * (o.prim_to_cons(x) - cn))

    eps = 1e-6
    fd = (F(w + eps * v) - F(w - eps * v)) / (2 * eps)
    assert C.rel_l2(got, fd) <= 1e-7


def test_ifunction_impose_ic_and_error():
    cfg = dict(C.vortex_cfg(4, 3, 3), cond_sys=1)
    o = Oracle(cfg)
    o0 = Oracle(dict(cfg, cond_sys=0))
    prm = [1.0, 1.0, 5.0, 1.0, 10.0]
    c = o0.impose_ic(1, prm)
    w = o.impose_ic(1, prm)
    np.testing.assert_allclose(w, o.cons_to_prim(c), rtol=0, atol=0)
    udot = np.sin(np.arange(w.size))
    np.testing.assert_allclose(o.ifunction(w, udot), udot + o0.residual(c), rtol=0, atol=1e-13)
    e0 = o0.l2_error(1, prm, c * 1.01)
    e1 = o.l2_error(1, prm, o.cons_to_prim(c * 1.01))
    np.testing.assert_allclose(e1, e0, rtol=1e-12)


def test_sod1d_maps():
    cfg = dict(C.sod_cfg(12, 3), cond_sys=1)
    o = Oracle(cfg)
    o0 = Oracle(dict(cfg, cond_sys=0))
    prm = [1.0, 0.0, 1.0, 0.125, 0.0, 0.1]
    c = o0.impose_ic(3, prm)
    w = o.impose_ic(3, prm)
    st = w.reshape(12, 3, 4)
    np.testing.assert_allclose(np.exp(st[:, 0]), c.reshape(12, 3, 4)[:, 0], rtol=1e-15)
    np.testing.assert_allclose(o.prim_to_cons(w), c, rtol=1e-14, atol=1e-15)

"""GPU parity of the implicit path with the positivity-preserving change of
variables (-cond_sys: state (log rho, u, [v,] log e); fun_implicit
petsc_solver_impl.hpp:626-655, complex-step MatMult :297-326, :386-401,
TSSetTransientVariable :572-575) against the CPU oracle."""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu


def _mods():
    import torch

    import fr_sdg_amd as fr
    from pyoracle import Oracle
    return torch, fr, Oracle


def _dev(torch, a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


CASES = {
    "vortex_p4": C.vortex_cfg(6, 5, 4),
    "explosion_hllc_vdw_p2": C.explosion_cfg(7, 4, 2, gas=1, riemann=1),
    "warped_vdw_p7": C.warped_cfg(3, 3, 7, filt=0),
    "bump_p3_ref_bc": dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1),
    "sod1d_p3": C.sod_cfg(20, 3),
}


def setup(name, seed=5):
    torch, fr, Oracle = _mods()
    cfg = dict(CASES[name], cond_sys=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    w = o.cons_to_prim(C.random_state(cfg, seed))
    return torch, ev, o, cfg, w


@pytest.mark.parametrize("name", sorted(CASES))
def test_ifunction_and_products(name):
    torch, ev, o, cfg, w = setup(name)
    print(f"[logvars] {name}", flush=True)
    rng = np.random.default_rng(2)
    udot = rng.standard_normal(w.size)
    v = rng.standard_normal(w.size)
    wd, vd = _dev(torch, w), _dev(torch, v)
    F = _host(ev.fun_implicit(wd, _dev(torch, udot)))
    assert C.rel_l2(F, o.ifunction(w, udot)) <= 1e-12
    ref = o.jvp_cstep(w, v, sigma=2.0)
    assert C.rel_l2(_host(ev.jvp_cstep(wd, vd, 2.0)), ref) <= 1e-11
    assert C.rel_l2(_host(ev.jvp(wd, vd, 2.0)), ref) <= 1e-9
    # adjoint through the map: <g, J v> = <J^T g, v>
    g = _dev(torch, rng.standard_normal(w.size))
    jv = ev.jvp(wd, vd, 2.0)
    jtg = ev.vjp(wd, g, 2.0)
    a, b = float(torch.dot(g, jv)), float(torch.dot(jtg, vd))
    assert abs(a - b) <= 1e-12 * float(torch.linalg.norm(g) * torch.linalg.norm(jv)), (a, b)


def test_impose_ic_and_error():
    torch, fr, Oracle = _mods()
    cfg = dict(C.vortex_cfg(6, 5, 4), cond_sys=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    prm = [1.0, 1.0, 5.0, 1.0, 10.0]
    w = ev.impose_ic(fr.IC_VORTEX, prm)
    wo = o.impose_ic(1, prm)
    assert C.rel_l2(_host(w), wo) <= 1e-14
    w.mul_(1.001)
    e = ev.compute_error(fr.IC_VORTEX, prm, w)
    np.testing.assert_allclose(e, o.l2_error(1, prm, _host(w)), rtol=1e-12)
    cfg1 = dict(C.sod_cfg(20, 3), cond_sys=1)
    w1 = fr.PhysicsEvaluator(cfg1).impose_ic(fr.IC_SOD, [1.0, 0.0, 1.0, 0.125, 0.0, 0.1])
    assert C.rel_l2(_host(w1), Oracle(cfg1).impose_ic(3, [1.0, 0.0, 1.0, 0.125, 0.0, 0.1])) <= 1e-14


def dense_beuler_logvars(o, w_n, dt, tol=1e-12, maxit=30):
    """Backward Euler on F(W) = (C(W) - C(W_n))/dt + R(C(W)) by dense Newton
    with the oracle's complex-step Jacobian through the map."""
    N = w_n.size
    cn = o.prim_to_cons(w_n)
    w = w_n.copy()
    for _ in range(maxit):
        F = o.ifunction(w, (o.prim_to_cons(w) - cn) / dt)
        if np.linalg.norm(F) <= tol * max(1.0, np.linalg.norm(cn) / dt):
            break
        J = np.zeros((N, N))
        for i in range(N):
            e = np.zeros(N)
            e[i] = 1.0
            J[:, i] = o.jvp_cstep(w, e, sigma=1.0 / dt)
        w = w - np.linalg.solve(J, F)
    return w


@pytest.mark.parametrize("jac_type,precond", [(0, 0), (1, 1), (0, 1), (0, 2), (0, 3), (1, 4)])
def test_beuler_with_transient_variables(jac_type, precond):
    """One implicit step on the tiny bump in log variables: JFNK == dense Newton,
    and C(W_{n+1}) == the conservative-variable step (the same discrete system)."""
    torch, fr, Oracle = _mods()
    base = dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1)
    cfg = dict(base, cond_sys=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u0 = Oracle(base).impose_ic(0, [1.0, 0.5 * 1.4 ** 0.5, 0.0, 1.0])
    u0 = u0 * (1.0 + 0.02 * np.sin(0.37 * np.arange(u0.size)))
    w0 = o.cons_to_prim(u0)
    dt = 0.05
    ref = dense_beuler_logvars(o, w0, dt)
    w = _dev(torch, w0)
    st = ev.beuler_step(w, dt, jac_type=jac_type, precond=precond, snes_rtol=1e-10, ksp_rtol=1e-8)
    print("[logvars newton]", jac_type, precond, st, flush=True)
    assert st["converged"]
    assert C.rel_l2(_host(w), ref) <= 1e-9
    # the conservative step solves the same equations
    ev0 = fr.PhysicsEvaluator(base)
    u = _dev(torch, u0)
    st0 = ev0.beuler_step(u, dt, jac_type=jac_type, precond=precond, snes_rtol=1e-10, ksp_rtol=1e-8)
    assert st0["converged"]
    assert C.rel_l2(o.prim_to_cons(_host(w)), _host(u)) <= 1e-9

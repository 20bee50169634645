"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bar: relative L2 <= 1e-12 for residual-type outputs (FP64), on seeded inputs
at sizes the oracle finishes in seconds.
"""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu

TOL = 1e-12


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


def smooth_1d(o, cfg):
    """rho = 1 + 0.2 sin(2 pi x), u = 0.1, p = 1 at the GLL nodes (a smooth 1D state)."""
    n = cfg["p"] + 1
    csi = o.tables()["csi"]
    h = cfg["lx"] / cfg["nx"]
    x = np.array([[h * (e + 0.5 * (1 + c)) for c in csi] for e in range(cfg["nx"])])
    rho = 1 + 0.2 * np.sin(2 * np.pi * x / cfg["lx"])
    st = np.stack([rho, 0.1 * rho, C.rhoE_of(cfg, rho, 0.01, 1.0)], axis=1)
    return st.reshape(-1)


RESIDUAL_CASES = {
    "vortex_p4": C.vortex_cfg(),
    "vortex_p4_1x1": C.vortex_cfg(1, 1, 4),
    "vortex_p1": C.vortex_cfg(5, 3, 1),
    "vortex_p8": C.vortex_cfg(3, 2, 8),
    "explosion_rusanov_ideal_p3": C.explosion_cfg(),
    "explosion_rusanov_vdw_p6": C.explosion_cfg(4, 3, 6, gas=1),
    "explosion_hllc_vdw_p2": C.explosion_cfg(7, 4, 2, gas=1, riemann=1),
    "bump_p5": C.bump_cfg(),
    "warped_vdw_p7": C.warped_cfg(),
    # curved meshes at low and odd orders: the even-odd line contractions with a
    # middle node (n = 3, 5), curved tiles of 64 and 10 elements, the LDS of the
    # largest tangent and adjoint tiles (n = 2)
    "warped_vdw_p1": C.warped_cfg(6, 5, 1),
    "bump_p2": C.bump_cfg(9, 4, 2),
    "warped_vdw_p4": C.warped_cfg(6, 4, 4),
    # curved and periodic, with x-ends whose face metrics differ (X from -1 to
    # 0.5): each copy of a wrap face carries its own metrics, so the adjoint
    # evaluates the outer copy's Riemann flux a second time
    "warped_vdw_p3_periodic": dict(C.warped_cfg(5, 4, 3), bc_type=C.PER, lx=1.5),
    "wall_p2_rusanov": C.wall_cfg(),
    "wall_p3_hllc": C.wall_cfg(3, 5, 3, 1),
    "vcjh_c_p3": dict(C.vortex_cfg(4, 4, 3), c_vcjh=0.01),
    "sod1d_p3_rusanov": C.sod_cfg(),
    "sod1d_p4_hllc_periodic": dict(C.sod_cfg(17, 4, 1), bc_type=[5, 5, 5, 5]),
    "sod1d_wall_p2": dict(C.sod_cfg(9, 2, 1), bc_type=[0, 0, 0, 0]),
}


@pytest.mark.parametrize("name", sorted(RESIDUAL_CASES))
def test_residual_parity(name):
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES[name]
    print(f"[parity] residual {name}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    ud = _dev(torch, u)
    R = _host(ev.residual_kernel(ud))
    Ro = o.residual(u)
    assert C.rel_l2(R, Ro) <= TOL, C.rel_l2(R, Ro)
    # rhs = -R ; F = udot + R (petsc_solver_impl.hpp:653,668)
    rhs = _host(ev.rhs_func(ud))
    assert np.array_equal(rhs, -R)
    udot = C.random_state(cfg, seed=7)
    F = _host(ev.fun_implicit(ud, _dev(torch, udot)))
    assert C.rel_l2(F, udot + Ro) <= TOL


@pytest.mark.parametrize("name", ["vortex_p4", "explosion_hllc_vdw_p2", "bump_p5", "warped_vdw_p7",
                                  "wall_p3_hllc", "sod1d_p3_rusanov"])
def test_complex_step_parity(name):
    """GPU complex-step J*v against the oracle's ComplexStepJacobianStrategy::matmult."""
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES[name]
    print(f"[parity] cstep {name}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    v = C.random_state(cfg, seed=11) - C.random_state(cfg, seed=12)
    sigma = 3.0
    got = _host(ev.jvp_cstep(_dev(torch, u), _dev(torch, v), sigma))
    ref = o.jvp_cstep(u, v, sigma)
    assert C.rel_l2(got, ref) <= 1e-11, C.rel_l2(got, ref)


@pytest.mark.parametrize("name", ["vortex_p4", "explosion_hllc_vdw_p2", "warped_vdw_p7", "wall_p3_hllc",
                                  "sod1d_p4_hllc_periodic", "warped_vdw_p1", "warped_vdw_p4",
                                  "warped_vdw_p3_periodic"])
def test_tangent_vs_complex_step(name):
    """Forward tangent J*v against the oracle's complex step (h = 1e-7, error O(h^2))."""
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES[name]
    print(f"[parity] tangent {name}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    v = C.random_state(cfg, seed=11) - C.random_state(cfg, seed=12)
    got = _host(ev.jvp(_dev(torch, u), _dev(torch, v), 0.5))
    ref = o.jvp_cstep(u, v, 0.5)
    assert C.rel_l2(got, ref) <= 1e-9, C.rel_l2(got, ref)


def test_tangent_bump_reference_mode():
    """With lambda BCs frozen (tangent_bc_mode 1) the tangent matches complex step."""
    torch, fr, Oracle = _mods()
    cfg = dict(C.bump_cfg(), tangent_bc_mode=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    v = C.random_state(cfg, seed=11) - C.random_state(cfg, seed=12)
    got = _host(ev.jvp(_dev(torch, u), _dev(torch, v)))
    ref = o.jvp_cstep(u, v)
    assert C.rel_l2(got, ref) <= 1e-9, C.rel_l2(got, ref)


def test_matmult_snapshot():
    torch, fr, Oracle = _mods()
    cfg = C.vortex_cfg()
    ev = fr.PhysicsEvaluator(cfg)
    u = _dev(torch, C.random_state(cfg))
    v = _dev(torch, C.random_state(cfg, seed=3))
    ev.pre_jv_eval(u, 2.5)
    a = _host(ev.matmult(v))
    b = _host(ev.jvp(u, v, 2.5))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["explosion_rusanov_vdw_p6", "warped_vdw_p7", "sod1d_p3_rusanov"])
def test_filter_parity(name):
    torch, fr, Oracle = _mods()
    cfg = dict(RESIDUAL_CASES[name], apply_filter=1, filt_cutoff=1, filt_strength=16, filt_alpha=36.0)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    got = _host(ev.limit_solution(_dev(torch, u)))
    assert C.rel_l2(got, o.filter(u)) <= TOL
    # in place (apply_stabilization)
    ud = _dev(torch, u)
    ev.apply_stabilization(ud)
    assert C.rel_l2(_host(ud), o.filter(u)) <= TOL


@pytest.mark.parametrize("name,stages", [("vortex_p4", 3), ("vortex_p4", 4), ("warped_vdw_p7", 3),
                                         ("explosion_rusanov_vdw_p6", 3), ("sod1d_p3_rusanov", 3),
                                         ("warped_vdw_p4", 3), ("warped_vdw_p1", 3), ("warped_vdw_p3_periodic", 3)])
def test_ssprk_parity(name, stages):
    torch, fr, Oracle = _mods()
    cfg = dict(RESIDUAL_CASES[name])
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    dt = 1e-3 if cfg["dims"] == 2 else 1e-4  # random nodal states: keep the 1D CFL small
    ud = _dev(torch, u)
    uo = u.copy()
    for k in range(3):
        ev.ssprk_step(ud, dt, stages=stages)
        uo = o.ssprk_step(uo, dt, stages=stages)
    assert C.rel_l2(_host(ud), uo) <= TOL
    # graph-captured run == stepped run (bitwise), on a smooth state (filtered
    # rough random states lose positivity within a few steps, as in the oracle)
    dt2 = 1e-4
    if cfg["dims"] == 2:
        u = o.impose_ic(4, [0.2, 0.3 * cfg["lx"], 0.5 * cfg["lx"], 0.5 * cfg["ly"], 1.0, 1.0])
        if cfg.get("geometry", 0) == 4:
            u = o.impose_ic(4, [0.2, 0.6, 0.0, 0.0, 1.0, 1.0])
    else:
        u = smooth_1d(o, cfg)
    ud2 = _dev(torch, u)
    ev.ssprk_run(ud2, dt2, 20, stages=stages)
    ud3 = _dev(torch, u)
    for k in range(20):
        ev.ssprk_step(ud3, dt2, stages=stages)
    a, b = _host(ud2), _host(ud3)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name,kind,prm", [
    ("vortex_p4", 1, [1.0, 1.0, 5.0, 1.0, 10.0]),
    ("explosion_rusanov_vdw_p6", 2, [1.0, 1.0, 0.125, 0.1, 0.4, 1.0, 1.0]),
    ("bump_p5", 0, [1.0, 0.5 * 1.4 ** 0.5, 0.0, 1.0]),
    ("sod1d_p3_rusanov", 3, [1.0, 0.0, 1.0, 0.125, 0.0, 0.1]),
])
def test_initial_condition_and_error(name, kind, prm):
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES[name]
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = _host(ev.impose_ic(kind, prm))
    uo = o.impose_ic(kind, prm)
    assert C.rel_l2(u, uo) <= 1e-14
    if cfg["dims"] == 2:
        pert = uo * (1.0 + 1e-3 * np.sin(np.arange(uo.size)))
        e = ev.compute_error(kind, prm, _dev(torch, pert), t=0.0)
        eo = o.l2_error(kind, prm, pert, t=0.0)
        assert C.rel_l2(e, eo) <= 1e-12


ADJ_CASES = ["vortex_p4", "vortex_p4_1x1", "vortex_p1", "vortex_p8", "explosion_rusanov_ideal_p3",
             "explosion_hllc_vdw_p2", "bump_p5", "warped_vdw_p7", "wall_p3_hllc", "vcjh_c_p3",
             "sod1d_p3_rusanov", "sod1d_p4_hllc_periodic", "sod1d_wall_p2", "warped_vdw_p1", "bump_p2",
             "warped_vdw_p4", "warped_vdw_p3_periodic"]


@pytest.mark.parametrize("name", ADJ_CASES)
def test_adjoint_dot_product(name):
    """<w, J v> = <J^T w, v> with J v from the tangent kernel (both exact linearisations)."""
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES[name]
    print(f"[parity] adjoint {name}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    u = _dev(torch, C.random_state(cfg))
    rng = np.random.default_rng(5)
    v = _dev(torch, rng.standard_normal(ev.ndofs))
    w = _dev(torch, rng.standard_normal(ev.ndofs))
    jv = ev.jvp(u, v, 0.7)
    jtw = ev.vjp(u, w, 0.7)
    a = float(torch.dot(w, jv))
    b = float(torch.dot(jtw, v))
    scale = float(torch.linalg.norm(w) * torch.linalg.norm(jv))
    assert abs(a - b) <= 1e-12 * scale, (a, b)


def test_adjoint_hllc_imaginary_sound_speed():
    """van der Waals states with c^2 < 0 at face nodes (rho 0.11, e 0.005: p < 0).
    The forward HLLC then compares NaN wave speeds, every test fails and F_R is
    taken, so residual and tangent stay finite; the adjoint must be finite too
    and satisfy <w, J v> = <J^T w, v> with the tangent kernel."""
    torch, fr, Oracle = _mods()
    cfg = RESIDUAL_CASES["explosion_hllc_vdw_p2"]
    n = cfg["p"] + 1
    st = C.random_state(cfg).reshape(-1, n * n, 4)
    rho, u, v, e = 0.11, 0.9, 0.25, 0.005
    a, b = cfg["vdw_a"], cfg["vdw_b"]
    p = 0.4 * rho * (e + a * rho) / (1 - b * rho) - a * rho * rho
    c2 = (1.4 * 0.4 * (e + a * rho) + 0.4 * a * rho * (2 - b * rho)) / (1 - b * rho) ** 2 - 2 * a * rho
    assert p < 0 and c2 < 0
    for el, node in ((3, n - 1), (9, 0), (17, n * n - 1), (20, n * (n - 1))):
        st[el, node] = [rho, rho * u, rho * v, rho * (e + 0.5 * (u * u + v * v))]
    ev = fr.PhysicsEvaluator(cfg)
    ud = _dev(torch, st.reshape(-1))
    rng = np.random.default_rng(6)
    vv = _dev(torch, rng.standard_normal(ev.ndofs))
    w = _dev(torch, rng.standard_normal(ev.ndofs))
    r = ev.residual_kernel(ud)
    jv = ev.jvp(ud, vv)
    jtw = ev.vjp(ud, w)
    assert bool(torch.isfinite(r).all()) and bool(torch.isfinite(jv).all())
    assert bool(torch.isfinite(jtw).all())
    lhs = float(torch.dot(w, jv))
    rhs = float(torch.dot(jtw, vv))
    assert abs(lhs - rhs) <= 1e-12 * float(torch.linalg.norm(w) * torch.linalg.norm(jv)), (lhs, rhs)


@pytest.mark.parametrize("name", ["vortex_p4", "bump_p5", "sod1d_p3_rusanov", "warped_vdw_p3_periodic"])
def test_adjoint_columns_vs_complex_step(name):
    """J^T e_i against rows of J built column by column from the oracle's complex step
    on a tiny mesh (reference-compatible BC mode, so complex step is exact)."""
    torch, fr, Oracle = _mods()
    base = RESIDUAL_CASES[name]
    cfg = dict(base, nx=min(base["nx"], 3), ny=min(base.get("ny", 1), 2), tangent_bc_mode=1)
    if cfg["dims"] == 1:
        cfg["nx"] = 4
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    N = ev.ndofs
    J = np.zeros((N, N))
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        J[:, i] = o.jvp_cstep(u, e)
    ud = _dev(torch, u)
    rng = np.random.default_rng(9)
    w = rng.standard_normal(N)
    got = _host(ev.vjp(ud, _dev(torch, w)))
    assert C.rel_l2(got, J.T @ w) <= 1e-9


def test_autograd_gradcheck():
    """FRResidual (forward = residual kernel, backward = adjoint kernel) under
    float64 gradcheck against finite differences, plus forward-mode (tangent)."""
    torch, fr, Oracle = _mods()
    cfg = C.vortex_cfg(2, 1, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u = _dev(torch, C.random_state(cfg)).requires_grad_(True)
    f = lambda x: fr.fr_residual(x, ev)  # noqa: E731
    assert torch.autograd.gradcheck(f, (u,), eps=1e-6, atol=1e-5, rtol=1e-4)


def test_sod_full_run_gpu_vs_oracle():
    """Config 1 full run (474 SSPRK(4,3) steps, filtered, HLLC) on the GPU
    against the oracle's run and the exact Riemann solution."""
    torch, fr, Oracle = _mods()
    import test_oracle_physics as TP
    cfg = dict(C.sod_cfg(riemann=1), **TP.SOD_FILTERED)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    prm = [1.0, 0.0, 1.0, 0.125, 0.0, 0.1]
    u = ev.impose_ic(fr.IC_SOD, prm)
    uo = o.impose_ic(3, prm)
    t = 0.0
    for dt in TP.sod_steps():
        ev.ssprk_step(u, dt, t=t, stages=4)
        uo = o.ssprk_step(uo, dt, t=t, stages=4)
        t += dt
    ug = _host(u)
    assert np.isfinite(ug).all()
    assert C.rel_l2(ug, uo) <= 1e-10
    eg, eo = TP.sod_l2_vs_exact(ug, cfg, t), TP.sod_l2_vs_exact(uo, cfg, t)
    assert abs(eg - eo) <= 1e-9 * eo and eg < 0.01


def test_vortex_run_gpu_vs_oracle():
    """Isentropic vortex, 8x8 p=3, 0.5 time units of SSP-RK3 on the GPU: state
    equal to the oracle's run and the same error against the exact solution."""
    torch, fr, Oracle = _mods()
    cfg = C.vortex_cfg(8, 8, 3)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    prm = [1.0, 1.0, 5.0, 1.0, 10.0]
    u = ev.impose_ic(fr.IC_VORTEX, prm)
    uo = o.impose_ic(1, prm)
    h = 10.0 / 8
    nst = int(np.ceil(0.5 / (0.1 * h / 3.0)))
    dt = 0.5 / nst
    ev.ssprk_run(u, dt, nst)
    for _ in range(nst):
        uo = o.ssprk_step(uo, dt)
    assert C.rel_l2(_host(u), uo) <= 1e-12
    e = ev.compute_error(fr.IC_VORTEX, prm, u, t=0.5)
    eo = o.l2_error(1, prm, uo, t=0.5)
    assert C.rel_l2(e, eo) <= 1e-9


def dense_beuler_oracle(o, u_n, dt, tol=1e-12, maxit=30):
    """Backward Euler by Newton with the dense Jacobian from the oracle's complex step."""
    N = u_n.size
    u = u_n.copy()
    for _ in range(maxit):
        F = (u - u_n) / dt + o.residual(u)
        if np.linalg.norm(F) <= tol * max(1.0, np.linalg.norm(u_n) / dt):
            break
        J = np.zeros((N, N))
        for i in range(N):
            e = np.zeros(N)
            e[i] = 1.0
            J[:, i] = o.jvp_cstep(u, e, sigma=1.0 / dt)
        u = u - np.linalg.solve(J, F)
    return u


@pytest.mark.parametrize("jac_type", [0, 1])
@pytest.mark.parametrize("precond", [0, 1, 2])
def test_beuler_jfnk_vs_dense_newton(jac_type, precond):
    """One implicit step on a tiny bump mesh (reference-compatible BCs so complex
    step and tangent linearise the same map) against a dense Newton solve."""
    torch, fr, Oracle = _mods()
    cfg = dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    g = 1.4
    u0 = o.impose_ic(0, [1.0, 0.5 * g ** 0.5, 0.0, 1.0])
    u0 = u0 * (1.0 + 0.02 * np.sin(0.37 * np.arange(u0.size)))  # away from the free-stream fixed point
    dt = 0.05
    ref = dense_beuler_oracle(o, u0, dt)
    u = _dev(torch, u0)
    st = ev.beuler_step(u, dt, jac_type=jac_type, precond=precond, snes_rtol=1e-10, ksp_rtol=1e-8)
    print("[newton]", jac_type, precond, st, flush=True)
    assert st["converged"]
    assert C.rel_l2(_host(u), ref) <= 1e-9


@pytest.mark.parametrize("stab", ["filter", "weno"])
def test_beuler_applies_post_step_stabilization(stab):
    """With the filter (or WENO limiter) on, the implicit step is the Newton
    solve followed by one stabilisation of the new state (TSSetPostStep /
    TSSetPostEvaluate, petsc_solver_impl.hpp:582-596)."""
    torch, fr, Oracle = _mods()
    extra = (dict(apply_filter=1, filt_cutoff=0, filt_strength=8, filt_alpha=36.0) if stab == "filter"
             else dict(apply_weno=1, shock_threshold=0.0))
    cfg = dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1, **extra)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u0 = o.impose_ic(0, [1.0, 0.5 * 1.4 ** 0.5, 0.0, 1.0])
    u0 = u0 * (1.0 + 0.02 * np.sin(0.37 * np.arange(u0.size)))
    dt = 0.05
    ref = o.filter(dense_beuler_oracle(o, u0, dt))
    assert C.rel_l2(ref, dense_beuler_oracle(o, u0, dt)) > 1e-6  # the stabilisation changes the state
    u = _dev(torch, u0)
    st = ev.beuler_step(u, dt, jac_type=0, precond=1, snes_rtol=1e-10, ksp_rtol=1e-8)
    print("[newton stab]", stab, st, flush=True)
    assert st["converged"]
    assert C.rel_l2(_host(u), ref) <= 1e-9


def test_batched_sweep_matches_per_instance_and_oracle():
    """8 instances drawn like the 64-instance sweep (mesh scaled down): the
    persistent batched stage kernels give bitwise the per-instance launches,
    and each instance matches the oracle's step."""
    torch, fr, Oracle = _mods()
    from fr_sdg_amd import configs as CF
    cfgs = CF.batch64(8, scale=1.0 / 16)
    cfgs[-1].update(gas_kind=1, vdw_a=0.1, vdw_b=0.02)
    b = fr.Batch(cfgs)
    u = b.empty()
    for k, iv in enumerate(b.instances):
        iv.impose_ic(fr.IC_EXPLOSION, CF.EXPLOSION_PRM, out=b.view(u, k))
        b.view(u, k).mul_(1.0 + 1e-3 * torch.sin(torch.arange(b.view(u, k).numel(), device="cuda",
                                                                dtype=torch.float64)))
    dts = [1e-4 * (1 + k) for k in range(len(cfgs))]
    b.set_dt(dts)
    u0 = _host(u).copy()
    ua = u.clone()
    ub = u.clone()
    for grid, div, mode in [(0, 1, 1), (0, 2, 1), (1024, 1, 1), (64, 2, 1), (0, 1, 2), (64, 1, 2), (0, 1, 0),
                            (7, 2, 0)]:
        b.set_tuning(grid, div, mode)
        ua.copy_(u)
        b.ssprk_step(ua)
        ub.copy_(u)
        b.ssprk_step_each(ub)
        # same tile code in a different kernel: equal up to FMA-contraction choices
        assert C.rel_l2(_host(ua), _host(ub)) <= 1e-14, (grid, div, mode)
    # the other tile shape of every order (fewest rim faces per element), in
    # the specialised, cache-chunked and all-orders launch modes
    for grid, mode in ((0, 1), (64, 2), (0, 0)):
        b.set_tuning(grid, 1, mode)
        for n in b.orders():
            b.set_tile_shape(n, 2)
        ua.copy_(u)
        b.ssprk_step(ua)
        assert C.rel_l2(_host(ua), _host(ub)) <= 1e-14, (grid, mode, "shape 2")
    with pytest.raises(fr.FrsdgError):
        b.set_tile_shape(3, 3)
    for k, cfg in enumerate(cfgs):
        o = Oracle(cfg)
        ref = o.ssprk_step(u0[b.offsets[k]:b.offsets[k + 1]], dts[k])
        assert C.rel_l2(_host(b.view(ua, k)), ref) <= TOL, k
    # graph replay == stepping, on smooth pulses (10 steps), persistent and
    # one-workgroup-per-tile launches
    for k, iv in enumerate(b.instances):
        iv.impose_ic(fr.IC_PULSE, [0.2, 0.5, 1.0, 1.0, 1.0, 1.0], out=b.view(u, k))
    for grid, mode in ((1024, 1), (0, 1), (0, 2)):
        b.set_tuning(grid, 1, mode)
        uc = u.clone()
        b.ssprk_run(uc, 10)
        ud = u.clone()
        for _ in range(10):
            b.ssprk_step(ud)
        assert np.isfinite(_host(uc)).all()
        assert np.array_equal(_host(uc), _host(ud)), (grid, mode)


def test_batched_tile_shapes_every_order():
    """Every order n = 3..9 on small ragged meshes (two per order, ideal and
    van der Waals): the three element tile shapes of the batched sweep
    (default, halved, fewest rim faces per element) against the per-instance
    launches, and against the oracle's step."""
    torch, fr, Oracle = _mods()
    from fr_sdg_amd import configs as CF
    cfgs = []
    for p in range(2, 9):
        cfgs.append(CF.explosion(7, p, vdw=False))
        cfgs.append(dict(CF.explosion(5, p, vdw=True), ny=6))
    b = fr.Batch(cfgs)
    u = b.empty()
    for k, iv in enumerate(b.instances):
        iv.impose_ic(fr.IC_EXPLOSION, CF.EXPLOSION_SMOOTH_PRM, out=b.view(u, k))
    dts = [2e-4 * (1 + (k % 3)) for k in range(len(cfgs))]
    b.set_dt(dts)
    u0 = _host(u).copy()
    ue = u.clone()
    b.ssprk_step_each(ue)
    ref = _host(ue)
    for shape in (0, 1, 2):
        print(f"[parity] batched tile shape {shape}", flush=True)
        b.set_tuning(0, 1, 1)
        for n in b.orders():
            b.set_tile_shape(n, shape)
        ua = u.clone()
        b.ssprk_step(ua)
        got = _host(ua)
        assert np.isfinite(got).all(), shape
        for k in range(len(cfgs)):
            sl = slice(b.offsets[k], b.offsets[k + 1])
            assert C.rel_l2(got[sl], ref[sl]) <= 1e-14, (shape, k, cfgs[k]["p"])
    for k, cfg in enumerate(cfgs):
        o = Oracle(cfg)
        r = o.ssprk_step(u0[b.offsets[k]:b.offsets[k + 1]], dts[k])
        assert C.rel_l2(ref[b.offsets[k]:b.offsets[k + 1]], r) <= TOL, (k, cfg["p"])

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
                                  "sod1d_p4_hllc_periodic"])
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
                                         ("explosion_rusanov_vdw_p6", 3), ("sod1d_p3_rusanov", 3)])
def test_ssprk_parity(name, stages):
    torch, fr, Oracle = _mods()
    cfg = dict(RESIDUAL_CASES[name])
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    dt = 1e-3
    ud = _dev(torch, u)
    uo = u.copy()
    for k in range(3):
        ev.ssprk_step(ud, dt, stages=stages)
        uo = o.ssprk_step(uo, dt, stages=stages)
    assert C.rel_l2(_host(ud), uo) <= TOL
    # graph-captured run == stepped run
    ud2 = _dev(torch, u)
    ev.ssprk_run(ud2, dt, 20, stages=stages)
    ud3 = _dev(torch, u)
    for k in range(20):
        ev.ssprk_step(ud3, dt, stages=stages)
    assert np.array_equal(_host(ud2), _host(ud3))


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

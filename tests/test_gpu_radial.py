"""GPU parity of the radial coordinate system's source terms (coord_sys
radial: physics_impl_2D.hpp:306-319, physics_impl_1D.hpp:70-90) against the
CPU oracle: residual, stage, tangent / complex-step products and the adjoint."""
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
    "axisym_p4_hllc": dict(dims=2, nx=7, ny=5, p=4, lx=2.0, ly=1.0, riemann=1, bc_type=[1, 1, 1, 0]),
    "axisym_p3_periodic_x": dict(dims=2, nx=6, ny=4, p=3, lx=2.0, ly=1.0, riemann=0, bc_type=[5, 5, 1, 1]),
    "axisym_warped_p7_vdw": dict(C.warped_cfg(4, 3, 7, filt=0), geom_param=[0.05, 0.0, 0.1, 0.0]),
    "radial1d_p3": dict(C.sod_cfg(24, 3), lx=1.0),
    "radial1d_p5_periodic": dict(C.sod_cfg(17, 5, 1), bc_type=C.PER),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_radial_residual_products_adjoint(name):
    torch, fr, Oracle = _mods()
    cfg = dict(CASES[name], coord_sys=1)
    print(f"[radial] {name}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    rng = np.random.default_rng(8)
    v = rng.standard_normal(u.size)
    ud, vd = _dev(torch, u), _dev(torch, v)
    R = o.residual(u)
    assert C.rel_l2(_host(ev.residual_kernel(ud)), R) <= 1e-12
    assert np.max(np.abs(R - Oracle(dict(cfg, coord_sys=0)).residual(u))) > 1e-3  # the source is there
    ref = o.jvp_cstep(u, v, sigma=1.5)
    assert C.rel_l2(_host(ev.jvp_cstep(ud, vd, 1.5)), ref) <= 1e-11
    assert C.rel_l2(_host(ev.jvp(ud, vd, 1.5)), ref) <= 1e-9
    g = _dev(torch, rng.standard_normal(u.size))
    jv = ev.jvp(ud, vd, 1.5)
    a, b = float(torch.dot(g, jv)), float(torch.dot(ev.vjp(ud, g, 1.5), vd))
    assert abs(a - b) <= 1e-12 * float(torch.linalg.norm(g) * torch.linalg.norm(jv)), (a, b)


@pytest.mark.parametrize("name", ["axisym_p4_hllc", "radial1d_p3"])
def test_radial_ssprk(name):
    torch, fr, Oracle = _mods()
    cfg = dict(CASES[name], coord_sys=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg)
    dt = 1e-4 if cfg["dims"] == 2 else 2e-5
    ud, uo = _dev(torch, u), u.copy()
    for _ in range(3):
        ev.ssprk_step(ud, dt)
        uo = o.ssprk_step(uo, dt)
    assert C.rel_l2(_host(ud), uo) <= 1e-12

"""Pins of the oracle's radial source terms (coord_sys radial:
physics_impl_2D.hpp:306-319, physics_impl_1D.hpp:70-90, get_r
physics_evaluator.hpp:148-155).

For a uniform state the volume and face terms vanish (free-stream
preservation, pinned elsewhere), so the residual is the source alone, in
closed form:
  2D: (1/r) (rho v, rho u v, rho v^2, v (E + p)) with r = y/2 at the mapped
      node (the G/r - p/r e_y form), zero for a fluid at rest;
  1D: (1/r) (rho u, rho u^2, u (E + p)) with r = x, on interior elements only
      (the reference adds it after evaluating the boundary elements).
"""
import numpy as np
import pytest

import cases as C
from pyoracle import Oracle


def uniform2d(o, rho, u, v, p, g=1.4):
    E = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    return np.tile([rho, rho * u, rho * v, E], o.ndofs // 4)


def node_y(cfg, o, warped=False):
    n = cfg["p"] + 1
    csi = o.tables()["csi"]
    nx, ny = cfg["nx"], cfg["ny"]
    j, i, l, k = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(n), np.arange(n), indexing="ij")
    cx = (1.0 / nx) * (i + 0.5 * (1.0 + csi[k]))
    cy = (1.0 / ny) * (j + 0.5 * (1.0 + csi[l]))
    if not warped:
        return (cfg["ly"] * cy).reshape(-1)
    A, x0, y0 = cfg["geom_param"][:3]
    X, Y = x0 + cfg["lx"] * cx, y0 + cfg["ly"] * cy
    return (Y + A * np.sin(np.pi * X) * np.sin(np.pi * Y)).reshape(-1)


def src2d(y, rho, u, v, p, g=1.4):
    r = 0.5 * y
    ir = np.where(r > 1e-8, 1.0 / np.where(r > 1e-8, r, 1.0), 0.0)
    E = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    s = np.stack([rho * v * ir, rho * u * v * ir, rho * v * v * ir, v * (E + p) * ir], axis=1)
    return s.reshape(-1)


@pytest.mark.parametrize("riemann", [0, 1])
def test_2d_rest_state_is_steady(riemann):
    cfg = dict(dims=2, nx=5, ny=4, p=3, lx=2.0, ly=1.0, riemann=riemann, bc_type=[5, 5, 1, 1], coord_sys=1)
    o = Oracle(cfg)
    R = o.residual(uniform2d(o, 1.3, 0.0, 0.0, 0.9))
    assert np.max(np.abs(R)) <= 1e-13


@pytest.mark.parametrize("warped", [False, True])
def test_2d_uniform_flow_source(warped):
    cfg = dict(dims=2, nx=5, ny=4, p=3, lx=2.0, ly=1.0, riemann=1, bc_type=[1, 1, 1, 1], coord_sys=1)
    if warped:
        cfg.update(geometry=4, geom_param=[0.05, 0.0, 0.2, 0.0])
    o = Oracle(cfg)
    rho, u, v, p = 0.9, 0.3, -0.2, 1.1
    R = o.residual(uniform2d(o, rho, u, v, p))
    want = src2d(node_y(cfg, o, warped), rho, u, v, p)
    assert C.rel_l2(R, want) <= 1e-12
    # Cartesian: same residual without the source is zero
    R0 = Oracle(dict(cfg, coord_sys=0)).residual(uniform2d(o, rho, u, v, p))
    assert np.max(np.abs(R0)) <= 1e-11


@pytest.mark.parametrize("periodic", [False, True])
def test_1d_uniform_flow_source(periodic):
    ne, p = 8, 3
    cfg = dict(C.sod_cfg(ne, p), lx=2.0, coord_sys=1)
    if periodic:
        cfg["bc_type"] = C.PER
    o = Oracle(cfg)
    n = p + 1
    rho, u, pr = 0.8, 0.4, 1.2
    E = pr / 0.4 + 0.5 * rho * u * u
    st = np.zeros((ne, 3, n))
    st[:, 0], st[:, 1], st[:, 2] = rho, rho * u, E
    R = o.residual(st.reshape(-1)).reshape(ne, 3, n)
    csi = o.tables()["csi"]
    h = 2.0 / ne
    r = (np.arange(ne)[:, None] + 0.5 * (csi[None, :] + 1.0)) * h
    ir = np.where(r > 1e-8, 1.0 / np.where(r > 1e-8, r, 1.0), 0.0)  # 0 on the axis
    want = np.stack([ir * rho * u, ir * rho * u * u, ir * u * (E + pr)], axis=1)
    if not periodic:
        want[0] = 0.0
        want[-1] = 0.0
    np.testing.assert_allclose(R, want, rtol=1e-12, atol=1e-11)

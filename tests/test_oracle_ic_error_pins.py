"""Oracle pins for the initial-state builder and the error norm (CPU).

* Gas::rhoE (gas_models.hpp:81-86, 144-157) is checked as the inverse of the
  pressure law that the fluxes use: the oracle's own flux of a state built by
  rhoE must carry back the prescribed pressure, for the ideal gas and for van
  der Waals, plus two values worked out by hand.
* The circular explosion (factory.hpp:400-465) is checked node by node against
  its definition: rho and p from the side of the circle r = 0.4 around the
  given centre the GLL node lies on, fluid at rest, on the Cartesian box and on
  the sine-warped map.
* compute_error (physics_impl_2D.hpp:203-265) is checked in absolute terms: an
  exact state offset by a constant delta in one variable has the error
  delta * sqrt(area) in that variable and 0 elsewhere.  A dropped detJ gives
  delta * sqrt(4E), a dropped GLL weight a node-count-dependent value: both fail.
"""
import math

import numpy as np
import pytest

from pyoracle import Oracle

GAMMA = 1.4
VDW = dict(gas_kind=1, vdw_a=0.1, vdw_b=0.02)


def _box(nx, ny, p, lx=2.0, ly=2.0, **kw):
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=lx, ly=ly, riemann=0, bc_type=[1, 1, 1, 1], **kw)


def _warped(n, p, **kw):
    return dict(dims=2, nx=n, ny=n, p=p, lx=2.0, ly=2.0, riemann=1, geometry=4,
                geom_param=[0.05, -1.0, -1.0, 0.0], bc_type=[1, 1, 1, 1], **kw)


def _pressure_from_flux(o, st):
    """p from the oracle's x-flux: F[1] = rho u^2 + p."""
    F, _ = o.flux2d(st)
    return F[1] - st[1] * st[1] / st[0]


@pytest.mark.parametrize("gas", ["ideal", "vdw"])
def test_rhoE_inverts_the_pressure_law(gas):
    cfg = _box(1, 1, 1, **(VDW if gas == "vdw" else {}))
    o = Oracle(cfg)
    rng = np.random.default_rng(5)
    for _ in range(50):
        rho, u, v, p = rng.uniform(0.2, 3.0), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.05, 5.0)
        st = o.impose_ic(0, [rho, u, v, p])[:4]
        assert st[0] == rho and abs(st[1] - rho * u) <= 1e-15 and abs(st[2] - rho * v) <= 1e-15
        assert abs(_pressure_from_flux(o, st) - p) <= 1e-13 * p


def test_rhoE_hand_values():
    # ideal: rho E = p / (gamma - 1) + rho |v|^2 / 2 = 1 / 0.4 + 0.5 * 2 * 0.25 = 2.75
    o = Oracle(_box(1, 1, 1))
    st = o.impose_ic(0, [2.0, 0.3, 0.4, 1.0])[:4]
    assert abs(st[3] - 2.75) <= 1e-15
    # van der Waals a = 0.1, b = 0.02 at rest, rho = 1, p = 1:
    # rho e = (p + a rho^2)(1 - b rho)/(gamma - 1) - a rho^2 = 1.1 * 0.98 / 0.4 - 0.1 = 2.595
    o = Oracle(_box(1, 1, 1, **VDW))
    st = o.impose_ic(0, [1.0, 0.0, 0.0, 1.0])[:4]
    assert abs(st[3] - 2.595) <= 1e-15


def _node_xy_box(cfg, csi):
    n, nx, ny = cfg["p"] + 1, cfg["nx"], cfg["ny"]
    hx, hy = cfg["lx"] / nx, cfg["ly"] / ny
    x = np.empty((ny, nx, n, n))
    y = np.empty_like(x)
    for j in range(ny):
        for i in range(nx):
            for l in range(n):
                for k in range(n):
                    x[j, i, l, k] = hx * (i + 0.5 * (1 + csi[k]))
                    y[j, i, l, k] = hy * (j + 0.5 * (1 + csi[l]))
    return x, y


def _node_xy_warped(cfg, csi):
    # x = X + A sin(pi X) sin(pi Y), y = Y + A sin(pi X) sin(pi Y), X, Y = -1 + 2 (element + (1 + csi)/2) / N
    X, Y = _node_xy_box(dict(cfg, lx=2.0, ly=2.0), csi)
    X, Y = X - 1.0, Y - 1.0
    s = 0.05 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    return X + s, Y + s


@pytest.mark.parametrize("case", ["box_ideal", "box_vdw", "warped_vdw"])
def test_explosion_ic_closed_form(case):
    if case == "warped_vdw":
        cfg = _warped(12, 5, **VDW)
        centre = (0.0, 0.0)
    else:
        cfg = _box(9, 7, 4, **(VDW if case == "box_vdw" else {}))
        centre = (1.0, 1.0)  # box length / 2 (factory.hpp:420-421)
    o = Oracle(cfg)
    csi = o.tables()["csi"]
    x, y = (_node_xy_warped if case == "warped_vdw" else _node_xy_box)(cfg, csi)
    prm = [1.0, 1.0, 0.125, 0.1, 0.4, centre[0], centre[1]]
    st = o.impose_ic(2, prm).reshape(x.shape + (4,))
    r = np.hypot(x - centre[0], y - centre[1])
    assert np.all(np.abs(r - 0.4) > 1e-9)  # no node on the circle
    inside = r < 0.4
    assert inside.any() and (~inside).any()
    np.testing.assert_array_equal(st[..., 0], np.where(inside, 1.0, 0.125))
    np.testing.assert_array_equal(st[..., 1], 0.0)
    np.testing.assert_array_equal(st[..., 2], 0.0)
    flat = st.reshape(-1, 4)
    p = np.array([_pressure_from_flux(o, s) for s in flat]).reshape(r.shape)
    np.testing.assert_allclose(p, np.where(inside, 1.0, 0.1), rtol=1e-13, atol=0)


@pytest.mark.parametrize("case", ["box", "warped"])
@pytest.mark.parametrize("var", [0, 3])
def test_l2_error_absolute_normalisation(case, var):
    cfg = _box(6, 5, 3) if case == "box" else _warped(8, 4, **VDW)
    o = Oracle(cfg)
    prm = [1.3, 0.2, -0.1, 0.9]
    u = o.impose_ic(0, prm)
    delta = 1e-3
    pert = u.reshape(-1, 4).copy()
    pert[:, var] += delta
    err = o.l2_error(0, prm, pert.reshape(-1))
    area = 4.0  # both domains are the square of side 2 (the warp fixes the boundary)
    # the warped map's detJ is not a polynomial: GLL quadrature of the area is exact to ~1e-9 here
    tol = 1e-11 if case == "box" else 1e-8  # (u + delta) - u carries ~1e-13 of rounding
    assert abs(err[var] - delta * math.sqrt(area)) <= tol * delta * math.sqrt(area), (err, var)
    for q in range(4):
        if q != var:
            assert err[q] == 0.0
    # the exact state itself has zero error
    assert np.all(o.l2_error(0, prm, u) == 0.0)

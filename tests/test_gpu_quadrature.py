"""Quadrature identities of the functionals' weights (frsdg_mass: w_k w_l detJ
per node; frsdg_boundary_weights: w ds n per boundary face node):

* the mass integrates x^a y^b exactly for a, b <= 2p - 1 on affine elements
  (GLL quadrature is exact to degree 2n - 3) and sums to the domain area on
  curved maps: the warped square (area 4) and the bump channel
  (l_x l_y - A w / 2 for the sin^2 bump of amplitude A and width w);
* the normals are the geometry model's face normals, +xi on x-faces and
  +eta on y-faces (outward on the right and top sides, inward on the left
  and bottom ones); the weighted normals of one side sum to the side's
  length times that direction (a constant field's flux), on the box, the
  warped map and the bump channel; on the bump's curved wall
  n ds = (-f', 1) dx, so sum w ds n_y = l_x whatever the bump's shape;
* the divergence theorem for the field (x, 0) closes the bump channel: the
  outward flux of x n_x w ds over the four sides equals the mass's area.
"""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu


def _mods():
    import torch

    import fr_sdg_amd as fr
    return torch, fr


def test_mass_exact_for_polynomials_on_affine_elements():
    torch, fr = _mods()
    cfg = C.vortex_cfg(5, 4, 4)  # [0, 10]^2
    ev = fr.PhysicsEvaluator(cfg)
    m = ev.mass().cpu().numpy().reshape(-1, 4)
    assert np.all(m == m[:, :1])  # the same weight for every equation of a node
    w = m[:, 0]
    csi = ev.tables()["csi"]
    n, nx, ny = cfg["p"] + 1, cfg["nx"], cfg["ny"]
    j, i, l, k = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(n), np.arange(n), indexing="ij")
    x = (10.0 / nx * (i + 0.5 * (1 + csi[k]))).reshape(-1)
    y = (10.0 / ny * (j + 0.5 * (1 + csi[l]))).reshape(-1)
    P = cfg["p"]
    for a in range(0, 2 * P):
        for b in (0, 1, 2 * P - 1):
            exact = 10.0 ** (a + 1) / (a + 1) * 10.0 ** (b + 1) / (b + 1)
            np.testing.assert_allclose((w * x ** a * y ** b).sum(), exact, rtol=1e-12)


@pytest.mark.parametrize("name", ["box", "warped", "bump"])
def test_area_and_boundary_normals(name):
    torch, fr = _mods()
    from fr_sdg_amd import configs as CF
    cfg = {"box": C.vortex_cfg(5, 4, 4), "warped": C.warped_cfg(6, 5, 5, filt=0), "bump": CF.bump(24, 8, 4)}[name]
    ev = fr.PhysicsEvaluator(cfg)
    lx, ly = cfg["lx"], cfg["ly"]
    w = ev.mass().cpu().numpy()[0::4]
    sides = {s: ev.boundary_weights(s).cpu().numpy().reshape(-1, 2) for s in (fr.LEFT, fr.RIGHT, fr.TOP, fr.BOTTOM)}
    sums = {s: v.sum(axis=0) for s, v in sides.items()}
    # left and right ends of the bump channel are straight (the bump is zero there)
    np.testing.assert_allclose(sums[fr.LEFT], [ly, 0.0], atol=1e-12 * ly)
    np.testing.assert_allclose(sums[fr.RIGHT], [ly, 0.0], atol=1e-12 * ly)
    np.testing.assert_allclose(sums[fr.TOP], [0.0, lx], atol=1e-12 * lx)
    np.testing.assert_allclose(sums[fr.BOTTOM][1], lx, rtol=1e-12)
    if name == "box":
        np.testing.assert_allclose(w.sum(), lx * ly, rtol=1e-12)
        np.testing.assert_allclose(sums[fr.BOTTOM][0], 0.0, atol=1e-12 * lx)
    elif name == "warped":
        np.testing.assert_allclose(w.sum(), lx * ly, rtol=1e-7)  # detJ is a trigonometric field: spectral accuracy
        np.testing.assert_allclose(sums[fr.BOTTOM][0], 0.0, atol=1e-12 * lx)
    else:
        A, x0, wb = cfg["geom_param"][0], cfg["geom_param"][1], cfg["geom_param"][3]
        area = lx * ly - 0.5 * A * wb
        np.testing.assert_allclose(w.sum(), area, rtol=1e-9)
        np.testing.assert_allclose(sums[fr.BOTTOM][0], 0.0, atol=1e-9)  # int f' dx = f(end) - f(start) = 0
        csi = ev.tables()["csi"]
        n, nx = cfg["p"] + 1, cfg["nx"]
        xb = (x0 + lx / nx * (np.arange(nx)[:, None] + 0.5 * (1 + csi[None, :]))).reshape(-1)
        # outward = -n on the left and bottom sides, +n on the right and top ones
        closure = -(x0 * sides[fr.LEFT][:, 0]).sum() + ((x0 + lx) * sides[fr.RIGHT][:, 0]).sum() \
            + (xb * sides[fr.TOP][:, 0]).sum() - (xb * sides[fr.BOTTOM][:, 0]).sum()
        np.testing.assert_allclose(closure, w.sum(), rtol=1e-9)

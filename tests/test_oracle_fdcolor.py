"""Pins of the oracle's preconditioning matrix and ILU(0) (oracle/fdcolor.py:
DMDA box colouring + MatFDColoringApply as the reference drives them,
petsc_solver_impl.hpp:72-95, 126-160; PETSc's ILU(0) PC).

* colouring: elements of one class are three apart in every direction (with
  the periodic wrap), so a block row meets at most one of them;
* differencing step: the perturbed states F is called at move each column
  once, by sqrt(1 + ||x||_2) sqrt(eps) ("wp") or sqrt(eps) x_i with |x_i|
  raised to 100 sqrt(eps) ("ds");
* a linear block-sparse F is recovered to FD accuracy (exactly with unit
  colour products), and on the oracle's implicit function the coloured
  matrix equals central differences of F on the box pattern, with the
  diagonal-neighbour blocks exactly zero; unit-colour complex-step products
  give the dense complex-step Jacobian (which on the bump differs from dF/du:
  the complex step freezes its lambda-bound boundary data);
* ILU(0): L and U vanish outside the pattern and (L U)_ij = A_ij on it (the
  defining property); with no fill possible (block tridiagonal, one element)
  it is the exact LU.  A full LU (pattern ignored) fails the first check.
"""
import numpy as np
import pytest

import cases as C
from fdcolor import EPS, UMIN, blocks_to_dense, box_neighbours, color_matrix, dense_to_blocks, ilu0, lu_solve
from pyoracle import Oracle


@pytest.mark.parametrize("dims,nx,ny,px,py", [(2, 7, 5, 0, 0), (2, 6, 9, 1, 1), (2, 3, 3, 1, 0), (1, 10, 1, 0, 0),
                                              (1, 9, 1, 1, 0)])
def test_colour_classes_separate_rows(dims, nx, ny, px, py):
    nbr, cls = box_neighbours(dims, nx, ny, px, py)
    for e in range(nbr.shape[0]):
        ms = nbr[e][nbr[e] >= 0]
        assert len(set(ms.tolist())) == len(ms)
        assert len(set(cls[ms].tolist())) == len(ms)  # one neighbour per class


def test_periodic_needs_multiple_of_three():
    with pytest.raises(ValueError):
        box_neighbours(2, 4, 3, 1, 0)
    with pytest.raises(ValueError):
        box_neighbours(1, 8, 1, 1, 0)


@pytest.mark.parametrize("method", [0, 1])
def test_difference_step(method):
    """The perturbations F sees are the colour's columns moved by the step:
    sqrt(1 + ||x||_2) sqrt(eps) ("wp"), sqrt(eps) x_i with |x_i| raised to
    100 sqrt(eps) ("ds"); every other entry is x itself."""
    rng = np.random.default_rng(0)
    dims, nx, ny, nd = 2, 4, 3, 5
    nbr, cls = box_neighbours(dims, nx, ny)
    x = rng.uniform(-2, 2, nx * ny * nd)
    x[3] = 1e-9  # below umin
    x[7] = -2e-8
    seen = []

    def F(v):
        seen.append(v.copy())
        return v * v

    color_matrix(F, x, nbr, cls, nd, method)
    assert np.array_equal(seen[0], x)  # w1 = F(x)
    if method == 0:
        want = np.full_like(x, np.sqrt(1 + np.linalg.norm(x)) * EPS)
    else:
        want = np.where(np.abs(x) < UMIN, np.where(x >= 0, UMIN, -UMIN), x) * EPS
    moved = np.zeros(x.size, dtype=int)
    for v in seen[1:]:
        d = v - x
        on = d != 0.0
        moved += on
        np.testing.assert_allclose(d[on], want[on], rtol=1e-7)
        assert len(set((np.nonzero(on)[0] // nd).tolist())) == on.sum()  # one column per element
    assert np.all(moved == 1)  # every column probed exactly once
    assert len(seen) == 1 + 9 * nd


def test_linear_map_recovered():
    rng = np.random.default_rng(1)
    dims, nx, ny, nd = 2, 5, 4, 3
    nbr, cls = box_neighbours(dims, nx, ny)
    B = rng.standard_normal((nx * ny, 9, nd, nd))
    B[nbr < 0] = 0.0
    A, pat = blocks_to_dense(B, nbr)
    b = rng.standard_normal(A.shape[0])
    x = rng.standard_normal(A.shape[0])
    P = color_matrix(lambda v: A @ v + b, x, nbr, cls, nd, 0)
    assert C.rel_l2(P, B) <= 1e-6
    P2 = color_matrix(None, x, nbr, cls, nd, 2, jv=lambda v: A @ v)
    assert C.rel_l2(P2, B) <= 1e-15


@pytest.mark.parametrize("name", ["bump", "vortex_periodic", "sod1d", "bump_logvars"])
def test_oracle_implicit_function(name):
    cfg = {"bump": dict(C.bump_cfg(4, 2, 2), tangent_bc_mode=1),
           "vortex_periodic": C.vortex_cfg(3, 3, 2),
           "sod1d": C.sod_cfg(7, 3),
           "bump_logvars": dict(C.bump_cfg(3, 2, 2), tangent_bc_mode=1, cond_sys=1)}[name]
    o = Oracle(cfg)
    dims = cfg["dims"]
    nx, ny = cfg["nx"], cfg.get("ny", 1)
    per = cfg["bc_type"][0] == 5
    nbr, cls = box_neighbours(dims, nx, ny if dims == 2 else 1, per, per and dims == 2)
    E = nbr.shape[0]
    u = C.random_state(cfg, 3)
    if cfg.get("cond_sys"):
        u = o.cons_to_prim(u)
    nd = u.size // E
    sigma = 7.0
    T = o.prim_to_cons if cfg.get("cond_sys") else (lambda v: v)
    un = u * 0.98
    Tn = T(un)

    def F(x):
        return o.ifunction(x, sigma * (T(x) - Tn))

    # dF/du by central differences (the complex step freezes lambda-bound
    # boundary data, so on the bump it is not dF/du)
    h = 1e-6
    Jcd = np.stack([(F(u + h * e) - F(u - h * e)) / (2 * h) for e in np.eye(u.size)], axis=1)
    Jb = dense_to_blocks(Jcd, nbr, nd)
    _, pat = blocks_to_dense(Jb, nbr)
    assert not np.any(Jcd[~pat])  # the residual couples face neighbours only
    P = color_matrix(F, u, nbr, cls, nd, 0)
    assert C.rel_l2(P, Jb) <= 1e-6
    assert C.rel_l2(color_matrix(F, u, nbr, cls, nd, 1), Jb) <= 1e-6
    if dims == 2:
        corners = [0, 2, 6, 8]
        assert np.max(np.abs(P[:, corners])) == 0.0
    Jcs = np.stack([o.jvp_cstep(u, e, sigma=sigma) for e in np.eye(u.size)], axis=1)
    P2 = color_matrix(None, u, nbr, cls, nd, 2, jv=lambda v: o.jvp_cstep(u, v, sigma=sigma))
    assert C.rel_l2(P2, dense_to_blocks(Jcs, nbr, nd)) <= 1e-14
    if "bump" not in name:
        assert C.rel_l2(P2, Jb) <= 1e-7


def random_block_matrix(nbr, nd, seed=2, shift=None):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nbr.shape[0], nbr.shape[1], nd, nd))
    B[nbr < 0] = 0.0
    A, pat = blocks_to_dense(B, nbr)
    A += np.eye(A.shape[0]) * (shift if shift is not None else 3.0 * nd * 9)
    return A, pat


@pytest.mark.parametrize("dims,nx,ny,per", [(2, 5, 4, 0), (2, 6, 3, 1), (1, 8, 1, 0), (1, 9, 1, 1)])
def test_ilu0_defining_property(dims, nx, ny, per):
    nd = 3
    nbr, _ = box_neighbours(dims, nx, ny, per, per and dims == 2)
    A, pat = random_block_matrix(nbr, nd)
    LU = ilu0(A, pat)
    assert not np.any(LU[~pat])
    L = np.tril(LU, -1) + np.eye(A.shape[0])
    U = np.triu(LU)
    np.testing.assert_allclose((L @ U)[pat], A[pat], rtol=0, atol=1e-11 * np.abs(A).max())
    # planted mistake: full LU of the same matrix fills outside the pattern
    full = ilu0(A, np.ones_like(pat))
    if dims == 2:
        assert np.any(full[~pat] != 0.0)
    x = np.sin(np.arange(A.shape[0]))
    np.testing.assert_allclose(L @ (U @ lu_solve(LU, x)), x, rtol=0, atol=1e-12)


def test_ilu0_without_fill_is_exact_lu():
    nd = 4
    nbr, _ = box_neighbours(1, 7, 1, 0, 0)  # block tridiagonal: no fill
    A, pat = random_block_matrix(nbr, nd, shift=1.0)
    LU = ilu0(A, pat)
    L = np.tril(LU, -1) + np.eye(A.shape[0])
    np.testing.assert_allclose(L @ np.triu(LU), A, rtol=0, atol=1e-11 * np.abs(A).max())
    nb1, _ = box_neighbours(2, 1, 1)
    A1, p1 = random_block_matrix(nb1, 6, shift=2.0)
    LU1 = ilu0(A1, p1)
    np.testing.assert_allclose((np.tril(LU1, -1) + np.eye(6)) @ np.triu(LU1), A1, atol=1e-12)

"""Test infrastructure (the oracle; only tests import it): CPU restatement of
the reference's preconditioning matrix and of the ILU(0) PETSc applies to it.

The reference (petsc_solver_impl.hpp:72-95 setup_colored_preconditioner,
:126-160 FDJacobianStrategy, :221-237 pre_jv_eval) builds P with PETSc:
DMCreateMatrix on its DMDA (one grid point per element, dof = the element's
DOFs, DMDA_STENCIL_BOX, stencil width 1: petsc_solver_impl.hpp:496-516),
DMCreateColoring(IS_COLORING_GLOBAL) and MatFDColoringApply on the implicit
function of the current stage (ts_coloring_wrapper :55-68).  PETSc is not
part of the reference tree and the reference pins no version; what follows
restates PETSc's published algorithms:

* DMCreateColoring for a 2D DMDA, box stencil of width s: colour of
  component k at grid point (i, j) = k + dof * ((i mod (2s+1)) +
  (2s+1) * (j mod (2s+1))); 1D: k + dof * (i mod (2s+1)); a periodic
  direction must hold a multiple of 2s+1 points.
* MatFDColoringApply (AIJ, block size 1): w1 = F(x); per colour, w3 = x with
  every column of the colour moved by its step, w2 = F(w3) - w1, entry
  (row, col) = w2[row] / step(col) for the pattern rows of col.  Step "wp"
  (MatFDColoring's default): one step for all columns,
  1/dx with dx = 1 / (sqrt(1 + ||x||_2) * error_rel); "ds": per column
  x_col raised to +-umin in magnitude, times error_rel; error_rel =
  sqrt(machine eps), umin = 100 sqrt(machine eps).
* PCILU with zero levels of fill, natural ordering, no pivoting: the IKJ
  form of Gaussian elimination restricted to the sparsity pattern; a pivot
  of magnitude <= 1e-12 is a zero pivot (an error).
"""
import numpy as np

EPS = float(np.sqrt(np.finfo(np.float64).eps))  # MatFDColoring error_rel
UMIN = 100.0 * EPS                               # MatFDColoring umin
ZEROPIVOT = 1e-12                                # PCFactorSetZeroPivot default (double)


def box_neighbours(dims, nx, ny=1, per_x=False, per_y=False):
    """nbr[E, S] (element of each stencil slot, -1 none; slot (dj+1)*3+(di+1)
    in 2D, di+1 in 1D) and the colour class of every element."""
    if dims == 1:
        ny, per_y = 1, False
    if (per_x and nx % 3) or (per_y and ny % 3):
        raise ValueError("periodic direction needs a multiple of 3 elements (DMDA colouring)")
    S = 9 if dims == 2 else 3
    E = nx * ny
    nbr = -np.ones((E, S), dtype=np.int64)
    cls = np.zeros(E, dtype=np.int64)
    djs = (-1, 0, 1) if dims == 2 else (0,)
    for j in range(ny):
        for i in range(nx):
            e = j * nx + i
            cls[e] = (i % 3) + 3 * (j % 3) if dims == 2 else i % 3
            for dj in djs:
                for di in (-1, 0, 1):
                    ii, jj = i + di, j + dj
                    if not 0 <= ii < nx:
                        if not per_x:
                            continue
                        ii %= nx
                    if not 0 <= jj < ny:
                        if not per_y:
                            continue
                        jj %= ny
                    slot = (dj + 1) * 3 + (di + 1) if dims == 2 else di + 1
                    nbr[e, slot] = jj * nx + ii
    return nbr, cls


def color_matrix(F, x, nbr, cls, nd, method=0, jv=None):
    """P[E, S, nd(col), nd(row)]: coloured differences of F at x (method 0
    "wp", 1 "ds") or, method 2, columns jv(unit colour vector)."""
    E, S = nbr.shape
    x = np.asarray(x, dtype=np.float64)
    P = np.zeros((E, S, nd, nd))
    if method != 2:
        w1 = F(x)
        if method == 0:
            unorm = np.linalg.norm(x)
            dx = 1.0 / (np.sqrt(1.0 + unorm) * EPS)
        else:
            d = x.copy()
            small = np.abs(d) < UMIN
            d[small] = np.where(d[small] >= 0.0, UMIN, -UMIN)
            vscale = 1.0 / (d * EPS)
    for c in range(int(cls.max()) + 1):
        members = np.nonzero(cls == c)[0]
        if members.size == 0:
            continue
        for k in range(nd):
            cols = members * nd + k
            if method == 2:
                v = np.zeros_like(x)
                v[cols] = 1.0
                w2 = jv(v)
            else:
                w3 = x.copy()
                w3[cols] += (1.0 / dx) if method == 0 else 1.0 / vscale[cols]
                w2 = F(w3) - w1
            for R in range(E):
                for s in range(S):
                    m = nbr[R, s]
                    if m < 0 or cls[m] != c:
                        continue
                    rows = w2[R * nd:(R + 1) * nd]
                    if method == 0:
                        P[R, s, k] = rows * dx
                    elif method == 1:
                        P[R, s, k] = rows * vscale[m * nd + k]
                    else:
                        P[R, s, k] = rows
                    break
    return P


def blocks_to_dense(P, nbr):
    """Dense matrix and boolean pattern of a block matrix in the P layout."""
    E, S, nd, _ = P.shape
    N = E * nd
    A = np.zeros((N, N))
    pat = np.zeros((N, N), dtype=bool)
    for e in range(E):
        for s in range(S):
            m = nbr[e, s]
            if m < 0:
                continue
            A[e * nd:(e + 1) * nd, m * nd:(m + 1) * nd] = P[e, s].T
            pat[e * nd:(e + 1) * nd, m * nd:(m + 1) * nd] = True
    return A, pat


def dense_to_blocks(A, nbr, nd):
    E, S = nbr.shape
    P = np.zeros((E, S, nd, nd))
    for e in range(E):
        for s in range(S):
            m = nbr[e, s]
            if m >= 0:
                P[e, s] = A[e * nd:(e + 1) * nd, m * nd:(m + 1) * nd].T
    return P


def ilu0(A, pat):
    """ILU(0) in IKJ order on the pattern: returns LU (unit-lower L below the
    diagonal, U on and above), zero outside the pattern."""
    a = np.where(pat, A, 0.0).astype(np.float64)
    N = a.shape[0]
    for i in range(1, N):
        for k in np.nonzero(pat[i, :i])[0]:
            if abs(a[k, k]) <= ZEROPIVOT:
                raise ZeroDivisionError("zero pivot")
            a[i, k] /= a[k, k]
            m = pat[i, k + 1:] & pat[k, k + 1:]
            a[i, k + 1:][m] -= a[i, k] * a[k, k + 1:][m]
    if np.any(np.abs(np.diag(a)) <= ZEROPIVOT):
        raise ZeroDivisionError("zero pivot")
    return a


def lu_solve(LU, x):
    """(L U)^-1 x for the combined factors of ilu0."""
    N = LU.shape[0]
    y = np.array(x, dtype=np.float64)
    for i in range(N):
        y[i] -= LU[i, :i] @ y[:i]
    for i in range(N - 1, -1, -1):
        y[i] = (y[i] - LU[i, i + 1:] @ y[i + 1:]) / LU[i, i]
    return y

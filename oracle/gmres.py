"""Test infrastructure (the oracle; only tests import it): CPU restatement of
the device GMRES of the implicit path, for residual-reduction curves.

The reference solves each Newton step's linear system with PETSc's KSPGMRES
(petsc_solver_impl.hpp:168-267 registers the operator; PETSc itself is not in
the reference tree and no version is pinned).  What is restated here is the
published restarted GMRES(m) with right preconditioning, classical
Gram-Schmidt (one pass, PETSc's default refinement "never") and Givens
rotations, in the order the device code runs it:

  x~ = 0; r = b; beta0 = ||b||; hist = [beta0]
  per cycle: beta = ||r||; stop if beta <= rtol beta0 (or 0, or the
  iteration limit); v_0 = r / beta; g = beta e_0
    per iteration j: w = A M^-1 v_j; h = V^T w; w -= V h; h_{j+1,j} = ||w||;
      v_{j+1} = w / h_{j+1,j}; rotate column j of H, g; hist += |g_{j+1}|;
      stop if |g_{j+1}| <= rtol beta0
    y = H^-1 g; x~ += V y; r = b - A M^-1 x~
  x = M^-1 x~
"""
import numpy as np


def gmres(A, b, Minv=None, restart=30, rtol=1e-5, max_it=10000):
    """Returns (x, hist, its): A a dense matrix or a callable, Minv a callable
    (None: identity)."""
    mat = (lambda v: A @ v) if not callable(A) else A
    prec = Minv if Minv is not None else (lambda v: v)
    b = np.asarray(b, dtype=np.float64)
    N = b.size
    m = max(1, restart)
    xt = np.zeros(N)
    hist = []
    beta0 = None
    its = 0
    first = True
    while True:
        r = b.copy() if first else b - mat(prec(xt))
        first = False
        beta = np.linalg.norm(r)
        if beta0 is None:
            beta0 = beta
            hist.append(beta)
        if not beta > rtol * beta0 or beta == 0.0 or its >= max_it:
            break
        V = np.zeros((m + 1, N))
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        j, done = 0, False
        while j < m and its < max_it:
            its += 1
            w = mat(prec(V[j]))
            h = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h
            hn = np.linalg.norm(w)
            H[:j + 1, j] = h
            H[j + 1, j] = hn
            if hn > 0.0:
                V[j + 1] = w / hn
            for i in range(j):
                a, bb = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * a + sn[i] * bb, -sn[i] * a + cs[i] * bb
            a, bb = H[j, j], H[j + 1, j]
            den = np.hypot(a, bb)
            cs[j], sn[j] = (a / den, bb / den) if den > 0 else (1.0, 0.0)
            H[j, j], H[j + 1, j] = den, 0.0
            g[j + 1], g[j] = -sn[j] * g[j], cs[j] * g[j]
            hist.append(abs(g[j + 1]))
            j += 1
            if abs(g[j]) <= rtol * beta0 or hn == 0.0:
                done = True
                break
        y = np.zeros(j)
        for i in range(j - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:j] @ y[i + 1:j]) / H[i, i] if H[i, i] != 0.0 else 0.0
        xt = xt + V[:j].T @ y
        if done or its >= max_it:
            break
    return prec(xt), np.array(hist), its

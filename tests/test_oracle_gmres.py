"""Pins of the oracle's GMRES (oracle/gmres.py), by properties of the
mathematics rather than by retyping it:

* exact termination: a matrix with k distinct eigenvalues (any basis) is
  solved in k iterations (the residual estimate drops to rounding at k, not
  before);
* the residual estimates are true residuals: hist[j] == ||b - A x_j|| where
  x_j is the solution GMRES returns when stopped after j iterations;
* within a cycle the estimates never increase; restarts keep the history
  going from the true residual;
* right preconditioning by the exact inverse converges in one iteration, and
  the returned x solves A x = b (np.linalg.solve) to the tolerance.
A restart residual formed without the preconditioner (b - A x~ instead of
b - A M^-1 x~) breaks the restart check below.
"""
import numpy as np
import pytest

from gmres import gmres


def rand_nonsym(n, seed, shift=4.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) / np.sqrt(n) + shift * np.eye(n), rng.standard_normal(n)


@pytest.mark.parametrize("k", [3, 7])
def test_exact_termination(k):
    rng = np.random.default_rng(k)
    n = 40
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    eig = np.repeat(np.linspace(1.0, 5.0, k), n // k + 1)[:n]
    A = Q @ np.diag(eig) @ Q.T
    b = rng.standard_normal(n)
    x, hist, its = gmres(A, b, restart=50, rtol=1e-12, max_it=50)
    assert its == k
    assert hist[k] <= 1e-10 * hist[0] and hist[k - 1] > 1e-6 * hist[0]
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


def test_estimates_are_true_residuals_and_monotone():
    A, b = rand_nonsym(60, 1, shift=1.2)
    _, hist, its = gmres(A, b, restart=100, rtol=0.0, max_it=25)
    assert its == 25 and hist.size == 26
    assert np.all(np.diff(hist) <= 1e-12 * hist[0])
    for j in (1, 5, 12, 25):
        xj, _, _ = gmres(A, b, restart=100, rtol=0.0, max_it=j)
        np.testing.assert_allclose(np.linalg.norm(b - A @ xj), hist[j], rtol=1e-8)


def test_restarts_and_preconditioning():
    A, b = rand_nonsym(80, 2, shift=1.5)
    Ainv = np.linalg.inv(A)
    x, hist, its = gmres(A, b, Minv=lambda v: Ainv @ v, restart=10, rtol=1e-12)
    assert its == 1 and hist[1] <= 1e-10 * hist[0]
    D = np.diag(1.0 / np.diag(A))
    x, hist, its = gmres(A, b, Minv=lambda v: D @ v, restart=8, rtol=1e-10, max_it=400)
    assert its > 8  # restarted at least once
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    # at every restart the estimate continues from a true residual of the
    # preconditioned iterate
    for j in (8, 16):
        xj, _, _ = gmres(A, b, Minv=lambda v: D @ v, restart=8, rtol=0.0, max_it=j)
        np.testing.assert_allclose(np.linalg.norm(b - A @ xj), hist[j], rtol=1e-8)

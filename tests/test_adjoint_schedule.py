"""CPU check of the reference binomial checkpointing schedule (tests/revolve_ref.py,
which frsdg_unsteady_adjoint must follow step for step) on a mock evaluator
whose step is a fixed nonlinear map: the gradient with few snapshots equals the
store-everything gradient, and the forward work obeys the revolve bound (at
most r forward sweeps of the T steps)."""
from math import comb

import numpy as np
import pytest
import torch

import revolve_ref as A


class MockEv:
    """u+ = tanh(M u) stepping; 'vjp' by the exact transpose."""
    cfg = {"apply_filter": 0, "dims": 2}

    def __init__(self, n=6, seed=0):
        g = torch.Generator().manual_seed(seed)
        self.M = torch.randn(n, n, generator=g, dtype=torch.float64) * 0.3

    # the UnsteadyAdjoint only uses these through adjoint_step / advance
    def ssprk_step(self, u, dt):
        u.copy_(torch.tanh(self.M @ u))


class MockAdjoint(A.Revolve):
    def adjoint_step(self, u, lam):
        self.adjoint_steps += 1
        y = torch.tanh(self.ev.M @ u)
        return self.ev.M.T @ ((1 - y * y) * lam)


@pytest.mark.parametrize("T,s", [(1, 0), (7, 0), (20, 1), (20, 3), (57, 2), (100, 5), (33, 32)])
def test_checkpointing_matches_store_all(T, s):
    ev = MockEv()
    u0 = torch.linspace(-1, 1, 6, dtype=torch.float64)
    lamT = torch.ones(6, dtype=torch.float64)
    ref = MockAdjoint(ev, 0.1, T, max_snapshots=T).gradient(u0, lamT)
    adj = MockAdjoint(ev, 0.1, T, max_snapshots=s)
    got = adj.gradient(u0, lamT)
    assert torch.allclose(got, ref, rtol=0, atol=1e-15)
    assert adj.adjoint_steps == T
    assert adj.peak_snapshots <= s
    if s > 0:
        r = A.repetitions(T, s)
        assert comb(s + r, s) >= T
        assert adj.forward_steps <= r * T


def test_gradient_vs_finite_difference_mock():
    ev = MockEv(seed=3)
    T = 15
    u0 = torch.linspace(-0.5, 0.7, 6, dtype=torch.float64)
    v = torch.cos(torch.arange(6, dtype=torch.float64))
    lamT = torch.ones(6, dtype=torch.float64)

    def J(u):
        x = u.clone()
        for _ in range(T):
            ev.ssprk_step(x, 0.1)
        return float(lamT @ x)
    g = MockAdjoint(ev, 0.1, T, max_snapshots=2).gradient(u0, lamT)
    eps = 1e-6
    fd = (J(u0 + eps * v) - J(u0 - eps * v)) / (2 * eps)
    assert abs(float(g @ v) - fd) <= 1e-8 * max(1.0, abs(fd))

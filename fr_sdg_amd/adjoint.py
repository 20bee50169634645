"""Unsteady discrete adjoint of the explicit FR time integration with
binomial checkpointing: a thin binding over frsdg_unsteady_adjoint (the
revolve schedule runs in the library's C++ host code, every state operation in
HIP kernels).

Forward map of one step (Shu-Osher SSPRK(3,3) with rhs = -R, then the
post-step filter F, as frsdg_ssprk_step):
    u1 = u + dt L(u),  u2 = 3/4 u + 1/4 (u1 + dt L(u1)),
    u3 = 1/3 u + 2/3 (u2 + dt L(u2)),  u+ = F u3,      L = -R.
Its transpose, with g = (dR/du)^T from the adjoint kernel:
    l3 = F^T l,  l2 = 2/3 l3 - 2/3 dt g(u2) l3,  l1 = 1/4 l2 - 1/4 dt g(u1) l2,
    l  <- 1/3 l3 + 3/4 l2 + l1 - dt g(u) l1.
"""
from math import comb

import torch


def repetitions(T, s):
    """Smallest r with C(s + r, s) >= T (forward sweeps of the binomial schedule)."""
    r = 0
    while comb(s + r, s) < T:
        r += 1
    return r


class UnsteadyAdjoint:
    """dJ/du_0 for J(u_T), u_T = S^T(u_0), with at most `max_snapshots` stored states."""

    def __init__(self, ev, dt, nsteps, max_snapshots=32):
        self.ev, self.dt, self.n = ev, dt, nsteps
        self.s = max(0, int(max_snapshots))
        self.forward_steps = 0   # forward steps recomputed by the last gradient()
        self.adjoint_steps = 0
        self.peak_snapshots = 0
        self.seconds = 0.0

    def run(self, u0):
        """u_T (u0 is not modified)."""
        u = u0.clone()
        self.ev.ssprk_run(u, self.dt, self.n)
        return u

    def gradient(self, u0, dJdu_T):
        """lambda_0 = (d u_T / d u_0)^T dJ/du_T; u0 is not modified."""
        lam, st = self.ev.unsteady_adjoint(u0, dJdu_T, self.dt, self.n, self.s)
        self.forward_steps = st["forward_steps"]
        self.adjoint_steps = st["adjoint_steps"]
        self.peak_snapshots = st["snapshots_used"]
        self.seconds = st["seconds"]
        return lam


def rho_squared(ev, u, m=None):
    """J = integral of rho^2 (GLL quadrature with the mapping Jacobian) and dJ/du."""
    m = ev.mass() if m is None else m
    if ev.cfg.get("dims", 2) == 2:
        rho = u.view(-1, 4)[:, 0]
        w = m.view(-1, 4)[:, 0]
        g = torch.zeros_like(u)
        g.view(-1, 4)[:, 0] = 2.0 * w * rho
    else:
        n = ev.n
        rho = u.view(-1, 3, n)[:, 0, :].reshape(-1)
        w = m.view(-1, 3, n)[:, 0, :].reshape(-1)
        g = torch.zeros_like(u)
        g.view(-1, 3, n)[:, 0, :] = (2.0 * w * rho).view(-1, n)
    return float(torch.dot(w, rho * rho)), g

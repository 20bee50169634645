"""Unsteady discrete adjoint of the explicit FR time integration with
binomial checkpointing.

Forward map of one step (Shu-Osher SSPRK(3,3) with rhs = -R, then the
post-step filter F, as frsdg_ssprk_step):
    u1 = u + dt L(u),  u2 = 3/4 u + 1/4 (u1 + dt L(u1)),
    u3 = 1/3 u + 2/3 (u2 + dt L(u2)),  u+ = F u3,      L = -R.
Its transpose, with g = (dR/du)^T from the adjoint kernel:
    l3 = F^T l,  l2 = 2/3 l3 - 2/3 dt g(u2) l3,  l1 = 1/4 l2 - 1/4 dt g(u1) l2,
    l  <- 1/3 l3 + 3/4 l2 + l1 - dt g(u) l1.
The stage states of a step are recomputed from the state at its start.

Checkpointing follows the binomial (revolve) schedule of Griewank: with s
free snapshot slots and r forward repetitions, L(s, r) = C(s+r, s) steps can
be reversed; a segment of T <= L(s, r) steps is split so that the later part
(reversed first, one slot fewer) has at most L(s-1, r) steps and the earlier
part at most L(s, r-1).  Snapshots live in device memory (one state each).
"""
from math import comb

import torch


def repetitions(T, s):
    """Smallest r with C(s + r, s) >= T."""
    r = 0
    while comb(s + r, s) < T:
        r += 1
    return r


class UnsteadyAdjoint:
    def __init__(self, ev, dt, nsteps, max_snapshots=32):
        self.ev, self.dt, self.n = ev, dt, nsteps
        self.s = max(0, int(max_snapshots))
        self.forward_steps = 0   # counts every forward step taken (incl. recomputation)
        self.adjoint_steps = 0
        self.peak_snapshots = 0
        self._live = 0
        self.on_adjoint_step = None  # optional callback(self) after each adjoint step

    # ---- forward ----
    def advance(self, u, k):
        for _ in range(k):
            self.ev.ssprk_step(u, self.dt)
        self.forward_steps += k
        return u

    def run(self, u0):
        u = u0.clone()
        return self.advance(u, self.n)

    # ---- one adjoint step at state u (start of the step) ----
    def adjoint_step(self, u, lam):
        ev, dt = self.ev, self.dt
        R = ev.residual_kernel(u)
        u1 = u - dt * R
        R1 = ev.residual_kernel(u1)
        u2 = 0.75 * u + 0.25 * (u1 - dt * R1)
        l3 = ev.filter_transpose(lam) if ev.cfg.get("apply_filter", 0) else lam
        l2 = (2.0 / 3.0) * (l3 - dt * ev.vjp(u2, l3))
        l1 = 0.25 * (l2 - dt * ev.vjp(u1, l2))
        self.adjoint_steps += 1
        out = (1.0 / 3.0) * l3 + 0.75 * l2 + l1 - dt * ev.vjp(u, l1)
        if self.on_adjoint_step is not None:
            self.on_adjoint_step(self)
        return out

    def _reverse(self, a, b, ua, lam, free):
        T = b - a
        if T == 1:
            return self.adjoint_step(ua, lam)
        if free == 0:
            for k in range(b - 1, a - 1, -1):
                x = ua.clone()
                self.advance(x, k - a)
                lam = self.adjoint_step(x, lam)
            return lam
        r = repetitions(T, free)
        second = min(T - 1, comb(free - 1 + r, free - 1))
        m = b - second
        x = ua.clone()
        self.advance(x, m - a)
        self._live += 1
        self.peak_snapshots = max(self.peak_snapshots, self._live)
        lam = self._reverse(m, b, x, lam, free - 1)
        del x
        self._live -= 1
        return self._reverse(a, m, ua, lam, free)

    def gradient(self, u0, dJdu_T):
        """lambda_0 = (d u_T / d u_0)^T dJ/du_T; u0 is not modified."""
        return self._reverse(0, self.n, u0, dJdu_T.clone(), self.s)


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

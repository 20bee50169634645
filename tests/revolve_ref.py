"""Reference binomial-checkpointing schedule (test infrastructure only): the
recursion of Griewank's revolve in plain Python over an evaluator that offers
ssprk_step (forward) and an adjoint_step(u, lam) hook.  The CPU tests pin it on
a mock map against store-everything gradients and the revolve work bound; the
GPU tests check that frsdg_unsteady_adjoint takes exactly the same number of
recomputed forward steps."""
from math import comb


def repetitions(T, s):
    """Smallest r with C(s + r, s) >= T."""
    r = 0
    while comb(s + r, s) < T:
        r += 1
    return r


class Revolve:
    def __init__(self, ev, dt, nsteps, max_snapshots=32):
        self.ev, self.dt, self.n = ev, dt, nsteps
        self.s = max(0, int(max_snapshots))
        self.forward_steps = 0   # every forward step taken during the reverse sweep
        self.adjoint_steps = 0
        self.peak_snapshots = 0
        self._live = 0

    def advance(self, u, k):
        for _ in range(k):
            self.ev.ssprk_step(u, self.dt)
        self.forward_steps += k
        return u

    def adjoint_step(self, u, lam):  # overridden by the tests
        raise NotImplementedError

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

    def gradient(self, u0, lamT):
        return self._reverse(0, self.n, u0, lamT.clone(), min(self.s, max(self.n - 1, 0)))

"""Pins of the oracle's SSP Runge-Kutta steps (which the device stage kernels
match in tests/test_gpu_parity.py), from the literature and the mathematics:

* one step equals the textbook composition with the residual, L = -R:
  SSPRK(3,3) (Shu & Osher 1988)
      u1 = u + dt L(u);  u2 = 3/4 u + 1/4 (u1 + dt L(u1));  u+ = 1/3 u + 2/3 (u2 + dt L(u2))
  SSPRK(4,3) (Kraaijevanger 1991; PETSc TSSSP rks3 with its default 4 stages)
      u1 = u + dt/2 L(u);  u2 = u1 + dt/2 L(u1);  u3 = 2/3 u + 1/3 (u2 + dt/2 L(u2));
      u+ = u3 + dt/2 L(u3);
* both are third order in time: on the smooth isentropic vortex the error
  against a fine-step solution falls by ~8 per halving of dt.
"""
import numpy as np
import pytest

import cases as C
from pyoracle import Oracle

VORTEX = [1.0, 1.0, 5.0, 1.0, 10.0]


def compose(o, u, dt, stages):
    L = lambda v: -o.residual(v)
    if stages == 3:
        u1 = u + dt * L(u)
        u2 = 0.75 * u + 0.25 * (u1 + dt * L(u1))
        return u / 3.0 + 2.0 / 3.0 * (u2 + dt * L(u2))
    u1 = u + 0.5 * dt * L(u)
    u2 = u1 + 0.5 * dt * L(u1)
    u3 = 2.0 / 3.0 * u + 1.0 / 3.0 * (u2 + 0.5 * dt * L(u2))
    return u3 + 0.5 * dt * L(u3)


@pytest.mark.parametrize("stages", [3, 4])
def test_step_is_the_textbook_composition(stages):
    cfg = C.vortex_cfg(4, 3, 3)
    o = Oracle(cfg)
    u = C.random_state(cfg, 9)
    dt = 2e-3
    np.testing.assert_allclose(o.ssprk_step(u, dt, stages=stages), compose(o, u, dt, stages), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("stages", [3, 4])
def test_third_order_in_time(stages):
    cfg = C.vortex_cfg(8, 8, 3)
    o = Oracle(cfg)
    u0 = o.impose_ic(1, VORTEX)
    T = 0.2  # 8 steps and more: inside the stability region of both schemes

    def run(nsteps):
        u = u0.copy()
        for _ in range(nsteps):
            u = o.ssprk_step(u, T / nsteps, stages=stages)
        return u

    ref = run(256)
    errs = [np.linalg.norm(run(n) - ref) for n in (8, 16, 32)]
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    assert all(6.5 <= r <= 9.5 for r in ratios), ratios

"""The oracle's 2D residual against a second, independent restatement in
vectorised numpy (tests/np_residual.py: Legendre-Vandermonde tables, HLLC in
Toro's standard star-state form, every face once) on small meshes: affine and
curved maps, all boundary kinds, both gas laws, both Riemann solvers, DG and a
VCJH c != 0 correction.  A dropped term, a transposed index or a wrong sign in
either restatement breaks the agreement."""
import numpy as np
import pytest

import cases as C
import np_residual as NR
from pyoracle import Oracle

TOL = 1e-12


def _cases():
    g = 1.4
    inflow = [1.0, 0.6, 0.1, 1.0 / (g - 1) + 0.5 * (0.36 + 0.01)]
    return {
        "vortex_periodic_hllc": C.vortex_cfg(3, 2, 4),
        "vortex_periodic_hllc_vcjh": dict(C.vortex_cfg(2, 3, 3), c_vcjh=0.35),
        "outflow_rusanov_vdw": C.explosion_cfg(3, 3, 3, gas=1, riemann=0),
        "walls_hllc": C.wall_cfg(3, 2, 2, riemann=1),
        "dirichlet_mixed_rusanov": dict(dims=2, nx=3, ny=2, p=3, lx=1.5, ly=1.0, riemann=0, bc_type=[2, 1, 2, 0],
                                        bc_state=[inflow, [0.0] * 4, inflow, [0.0] * 4]),
        "warped_outflow_hllc_vdw": C.warped_cfg(3, 3, 5, filt=0),
        "warped_slipwall_rusanov": dict(C.warped_cfg(2, 3, 4, filt=0), riemann=0, bc_type=[1, 1, 1, 6]),
        "warped_periodic_hllc": dict(C.warped_cfg(3, 2, 3, filt=0), bc_type=C.PER, gas_kind=0),
    }


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_residual_matches_numpy_restatement(name):
    cfg = CASES[name]
    o = Oracle(cfg)
    u = C.random_state(cfg)
    ref = NR.residual(cfg, u)
    got = o.residual(u)
    assert C.rel_l2(got, ref) <= TOL, (name, C.rel_l2(got, ref))


def test_numpy_tables_match_oracle_tables():
    for p in range(1, 9):
        for c in (0.0, 0.4):
            x, D, dL, dR = NR.tables(p, c)
            t = Oracle(dict(dims=2, nx=1, ny=1, p=p, lx=1.0, ly=1.0, c_vcjh=c, bc_type=C.PER)).tables()
            n = p + 1
            assert np.allclose(x, t["csi"], atol=1e-14, rtol=0)
            assert np.allclose(D, np.asarray(t["D"]).reshape(n, n), atol=1e-11 * p * p, rtol=0)
            assert np.allclose(dL, t["dL"], atol=1e-11 * p * p, rtol=0)
            assert np.allclose(dR, t["dR"], atol=1e-11 * p * p, rtol=0)


def test_planted_mistake_is_caught():
    """A sign slip in the numpy HLLC star state (a mistake of the kind this
    check is for) moves the residual far outside the tolerance."""
    cfg = CASES["walls_hllc"]
    u = C.random_state(cfg)
    good = NR.residual(cfg, u)
    orig = NR.riemann

    def bad(UL, UR, nx, ny, gas, kind):
        f = orig(UL, UR, nx, ny, gas, kind)
        return f * np.array([1.0, 1.0, 1.0, 1.0 + 1e-6])
    NR.riemann = bad
    try:
        assert C.rel_l2(NR.residual(cfg, u), good) > 1e-9
    finally:
        NR.riemann = orig

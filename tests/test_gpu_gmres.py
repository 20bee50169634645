"""Residual-reduction curves of the device GMRES (frsdg_linear_solve) against
the oracle's GMRES (oracle/gmres.py) on the same operator sigma I + J(u): the
dense Jacobian from the oracle's complex step (reference-compatible boundary
linearisation, so the tangent kernel linearises the same map), with no
preconditioner, the element block-Jacobi preconditioner (exact diagonal
blocks; the device keeps their inverses in FP32) and ILU(0) of the coloured
matrix (exact columns) restated by oracle/fdcolor.py."""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu


def _mods():
    import torch

    import fr_sdg_amd as fr
    from pyoracle import Oracle
    return torch, fr, Oracle


@pytest.mark.parametrize("precond,tol", [(0, 1e-9), (1, 1e-5), (4, 1e-9)])
def test_gmres_history_matches_oracle(precond, tol):
    from fdcolor import blocks_to_dense, box_neighbours, dense_to_blocks, ilu0, lu_solve
    from gmres import gmres
    torch, fr, Oracle = _mods()
    cfg = dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg, 7)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(u.size)
    sigma = 20.0
    A = np.stack([o.jvp_cstep(u, e, sigma=sigma) for e in np.eye(u.size)], axis=1)
    nbr, _ = box_neighbours(2, cfg["nx"], cfg["ny"])
    nd = u.size // nbr.shape[0]
    if precond == 0:
        Minv = None
    elif precond == 1:
        inv = [np.linalg.inv(A[e * nd:(e + 1) * nd, e * nd:(e + 1) * nd]) for e in range(nbr.shape[0])]
        Minv = lambda v: np.concatenate([inv[e] @ v[e * nd:(e + 1) * nd] for e in range(len(inv))])
    else:
        Ad, pat = blocks_to_dense(dense_to_blocks(A, nbr, nd), nbr)
        LU = ilu0(Ad, pat)
        Minv = lambda v: lu_solve(LU, v)
    opts = dict(ksp_restart=20, ksp_max_it=80, ksp_rtol=1e-10)
    xo, ho, its = gmres(A, b, Minv, restart=20, rtol=1e-10, max_it=80)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    x, st, hist = ev.linear_solve(dev(u), dev(b), sigma, jac_type=0, precond=precond, **opts)
    torch.cuda.synchronize()
    hist = np.array(hist)
    print(f"[gmres] precond {precond}: device {st['ksp_its']} its, oracle {its}; "
          f"reduction {hist[-1] / hist[0]:.2e} vs {ho[-1] / ho[0]:.2e}", flush=True)
    assert abs(st["ksp_its"] - its) <= (0 if precond != 1 else 1)
    assert hist.size == st["ksp_its"] + 1
    n = min(hist.size, ho.size)
    assert n >= 5
    assert np.max(np.abs(hist[:n] - ho[:n])) <= tol * ho[0]
    big = ho[:n] > 1e-6 * ho[0]
    assert np.max(np.abs(hist[:n][big] / ho[:n][big] - 1.0)) <= 1e3 * tol
    assert C.rel_l2(x.cpu().numpy(), xo) <= (1e-8 if precond != 1 else 1e-5)
    # the true final residual agrees with the oracle's
    r_o = np.linalg.norm(A @ xo - b)
    assert abs(st["fnorm"] - r_o) <= 1e-3 * r_o + 1e-12 * np.linalg.norm(b)


@pytest.mark.parametrize("precond", [3, 4])
def test_adjoint_solve_with_transposed_ilu(precond):
    """(sigma I + J)^T lambda = g by GMRES on the adjoint kernel, preconditioned
    by ILU(0) of the coloured matrix applied transposed, against a dense solve;
    far fewer iterations than without a preconditioner."""
    torch, fr, Oracle = _mods()
    cfg = dict(C.bump_cfg(4, 2, 3), tangent_bc_mode=1)
    ev = fr.PhysicsEvaluator(cfg)
    o = Oracle(cfg)
    u = C.random_state(cfg, 7)
    g = np.random.default_rng(5).standard_normal(u.size)
    sigma = 20.0
    A = np.stack([o.jvp_cstep(u, e, sigma=sigma) for e in np.eye(u.size)], axis=1)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    lam, st = ev.adjoint_solve(dev(u), dev(g), sigma, precond=precond, ksp_rtol=1e-12, ksp_restart=60,
                               ksp_max_it=600)
    _, st0 = ev.adjoint_solve(dev(u), dev(g), sigma, precond=0, ksp_rtol=1e-12, ksp_restart=60, ksp_max_it=600)
    print(f"[adjoint ilu] precond {precond}: {st['ksp_its']} its (none: {st0['ksp_its']})", flush=True)
    assert C.rel_l2(lam.cpu().numpy(), np.linalg.solve(A.T, g)) <= 1e-9
    assert st["ksp_its"] * 4 <= st0["ksp_its"]

"""GPU checks of the derivative-based solvers: the unsteady discrete adjoint
with binomial checkpointing against central finite differences, and the
steady adjoint (J^T lambda) solve of a wall-pressure functional against
central finite differences through an implicit step."""
import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu


def _mods():
    import torch

    import fr_sdg_amd as fr
    return torch, fr


def _dev(torch, a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("filt", [0, 1])
def test_unsteady_adjoint_rho2_vs_fd(filt):
    """Curved (warped) VdW problem, 12 SSP-RK3 steps (+ filter), J = int rho^2 at
    the end: <dJ/du0, v> from the adjoint with 2 snapshots against central FD;
    and the 2-snapshot gradient equals the store-all gradient."""
    torch, fr = _mods()
    from fr_sdg_amd import adjoint as A
    from fr_sdg_amd import configs as CF
    cfg = CF.warped(6, 3, filt=filt)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_PULSE, [0.3, 0.5, 0.1, -0.1, 1.0, 1.0])  # smooth pulse
    T, dt = 12, 2e-3
    m = ev.mass()

    def J(u):
        adj = A.UnsteadyAdjoint(ev, dt, T)
        return A.rho_squared(ev, adj.run(u), m)[0]
    adj = A.UnsteadyAdjoint(ev, dt, T, max_snapshots=2)
    uT = adj.run(u0)
    _, g = A.rho_squared(ev, uT, m)
    lam = adj.gradient(u0, g)
    full = A.UnsteadyAdjoint(ev, dt, T, max_snapshots=T).gradient(u0, g)
    assert float(torch.linalg.norm(lam - full) / torch.linalg.norm(full)) <= 1e-12
    v = torch.cos(torch.arange(u0.numel(), device="cuda", dtype=torch.float64)) * u0 * 1e-2
    eps = 1e-4
    fd = (J(u0 + eps * v) - J(u0 - eps * v)) / (2 * eps)
    ad = float(torch.dot(lam, v))
    print(f"[adjoint] filt={filt} adjoint {ad:.12e} fd {fd:.12e}", flush=True)
    # central-difference error budget: rounding ~ 1e-16 |J| / (eps |dJ.v|) ~ 1e-7 plus
    # truncation O(eps^2) of the same order
    assert abs(ad - fd) <= 3e-7 * abs(fd)


def wall_pressure(torch, ev, u, W):
    """J = sum over bottom-wall face nodes of (w ds n_y) p, and dJ/du."""
    cfg = ev.cfg
    n = cfg["p"] + 1
    g = cfg.get("gamma", 1.4)
    st = u.view(cfg["ny"], cfg["nx"], n * n, 4)[0, :, :n, :]  # elements j=0, nodes l=0 (node = k)
    rho, mu, mv, E = st[..., 0], st[..., 1], st[..., 2], st[..., 3]
    p = (g - 1) * (E - 0.5 * (mu * mu + mv * mv) / rho)
    wy = W[..., 1]
    Jv = float((wy * p).sum())
    grad = torch.zeros_like(u)
    gv = grad.view(cfg["ny"], cfg["nx"], n * n, 4)[0, :, :n, :]
    uu, vv = mu / rho, mv / rho
    gv[..., 0] = wy * (g - 1) * 0.5 * (uu * uu + vv * vv)
    gv[..., 1] = -wy * (g - 1) * uu
    gv[..., 2] = -wy * (g - 1) * vv
    gv[..., 3] = wy * (g - 1)
    return Jv, grad


@pytest.mark.parametrize("precond", [0, 1, 2, 3, 4])
def test_wall_pressure_adjoint_vs_fd(precond):
    """Bump channel, one backward-Euler step u1(u0): dJ(u1)/du0 = sigma lambda with
    (sigma I + J(u1))^T lambda = dJ/du1 (adjoint kernel + GMRES), against
    central differences with step 1e-7 through tightly converged implicit solves."""
    torch, fr = _mods()
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(4, 2, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    dt = 0.02
    tight = dict(snes_rtol=1e-15, snes_atol=1e-13, ksp_rtol=1e-13, ksp_restart=100, snes_max_it=30)
    W = ev.boundary_weights(fr.BOTTOM)

    def solve(u):
        x = u.clone()
        ev.beuler_step(x, dt, **tight)
        return x
    u1 = solve(u0)
    J1, g1 = wall_pressure(torch, ev, u1, W)
    lam, st = ev.adjoint_solve(u1, g1, 1.0 / dt, precond=precond, ksp_rtol=1e-13, ksp_restart=100)
    print("[adjoint solve]", st, flush=True)
    grad0 = lam / dt
    v = torch.sin(1.0 + torch.arange(u0.numel(), device="cuda", dtype=torch.float64))
    eps = 1e-7
    Jp = wall_pressure(torch, ev, solve(u0 + eps * v), W)[0]
    Jm = wall_pressure(torch, ev, solve(u0 - eps * v), W)[0]
    fd = (Jp - Jm) / (2 * eps)
    ad = float(torch.dot(grad0, v))
    print(f"[wall] adjoint {ad:.12e} fd {fd:.12e}", flush=True)
    assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-9


@pytest.mark.parametrize("from_initial", [True, False])
def test_pseudo_transient_step_rule_and_replay(from_initial):
    """frsdg_pseudo_run follows TSStep_Pseudo + TSPseudoTimeStepDefault (SER
    rule evaluated before each step from the current state): the first step
    takes inc * dt0, and step k the rule's size from steady-residual norms
    that this test computes itself on the states of a step-by-step replay
    (frsdg_beuler_step + frsdg_ifunction with U_t = 0), not from the run's
    returned norms; the replay equals the run bitwise and the steady residual
    falls (8x4 p=2 bump from the free stream)."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(8, 4, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    dt0 = 5.0 * h / CF.sound_speed_max(u0, cfg)
    opts = dict(jac_type=0, precond=1, ksp_restart=60, ksp_max_it=600, ksp_rtol=1e-4, snes_rtol=1e-6,
                snes_max_it=8)
    inc = 1.5
    u = u0.clone()
    res = ev.pseudo_run(u, dt0, inc, 6, from_initial_dt=from_initial, **opts)
    assert len(res["seconds"]) == 6 and all(b >= a for a, b in zip(res["seconds"], res["seconds"][1:]))
    dts = res["dt"]
    print("[pseudo]", from_initial, dts, res["fnorm"], flush=True)
    assert res["steps"] == 6
    assert dts[0] == inc * dt0  # f_0 / f_0 = 1 on either rule

    def steady(w, t):
        return float(torch.linalg.norm(ev.fun_implicit(w, torch.zeros_like(w), t=t)))

    # replay: the expected size of every step from independently computed norms
    w = u0.clone()
    t = 0.0
    f0 = steady(w, t)
    f_prev, dt_prev = f0, dt0
    for k in range(6):
        f = steady(w, t)
        want = inc * dt0 * f0 / f if from_initial else inc * dt_prev * f_prev / f
        assert dts[k] == pytest.approx(want, rel=1e-13), k
        ev.beuler_step(w, dts[k], t=t, **opts)
        t += dts[k]
        f_prev, dt_prev = f, dts[k]
        assert res["fnorm"][k] == pytest.approx(steady(w, t), rel=1e-13)
    torch.cuda.synchronize()
    assert torch.equal(w, u)
    assert res["fnorm"][-1] < f0
    # t_max: the last step lands on it
    v = u0.clone()
    r2 = ev.pseudo_run(v, dt0, inc, 50, from_initial_dt=from_initial, t_max=2.5 * dt0, **opts)
    assert abs(sum(r2["dt"]) - 2.5 * dt0) <= 1e-12 * dt0


def test_block_gauss_seidel_cuts_gmres_iterations():
    """Red-black element block Gauss-Seidel (precond 2) needs fewer GMRES
    iterations than block-Jacobi for the same backward-Euler solve, and gives
    the same converged step."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(16, 4, 3)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    u0.mul_(1.0 + 0.01 * torch.sin(0.37 * torch.arange(u0.numel(), device="cuda", dtype=torch.float64)))
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    dt = 2.0 * h / CF.sound_speed_max(u0, cfg)
    out, its = {}, {}
    for pc in (1, 2):
        u = u0.clone()
        st = ev.beuler_step(u, dt, jac_type=0, precond=pc, ksp_restart=100, ksp_max_it=2000, ksp_rtol=1e-8,
                            snes_rtol=1e-10, snes_max_it=10)
        print("[gs]", pc, st, flush=True)
        assert st["converged"]
        out[pc], its[pc] = u, st["ksp_its"]
    assert its[2] < its[1]
    assert float(torch.linalg.norm(out[2] - out[1]) / torch.linalg.norm(out[1])) <= 1e-8


def test_snes_step_norm_test_stops_first():
    """snes_stol: with a residual target out of reach, the Newton loop stops on
    ||lambda du|| <= stol ||u|| (SNES_CONVERGED_SNORM_RELATIVE, converged = 2)
    after fewer iterations than with the test off, on a state whose last step
    was already below the threshold."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(8, 4, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    dt = 2.0 * h / CF.sound_speed_max(u0, cfg)
    base = dict(jac_type=0, precond=1, ksp_restart=60, ksp_max_it=600, ksp_rtol=1e-10, snes_rtol=1e-30,
                snes_atol=0.0, snes_max_it=12)
    a = u0.clone()
    sa = ev.beuler_step(a, dt, snes_stol=0.0, **base)
    b = u0.clone()
    sb = ev.beuler_step(b, dt, snes_stol=1e-6, **base)
    print("[stol]", sa, sb, flush=True)
    assert sa["converged"] == 0 and sa["newton_its"] == 12
    assert sb["converged"] == 2 and 1 <= sb["newton_its"] < 12
    # the step-norm stop leaves a state close to the fully iterated one
    assert float(torch.linalg.norm(a - b) / torch.linalg.norm(a)) <= 1e-6


@pytest.mark.parametrize("precond", [1, 3])
def test_left_preconditioned_gmres_solves_the_same_system(precond):
    """ksp_pc_side 1 (PETSc's GMRES default: left preconditioning, test on the
    preconditioned residual) and 2 (flexible GMRES) solve the same linear
    system as the right-preconditioned solver; the left history starts at
    ||M^-1 b||, and the flexible solution's true residual is its estimate."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(8, 4, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    u.mul_(1.0 + 0.01 * torch.sin(0.37 * torch.arange(u.numel(), device="cuda", dtype=torch.float64)))
    b = torch.cos(torch.arange(u.numel(), device="cuda", dtype=torch.float64))
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    sigma = CF.sound_speed_max(u, cfg) / (4.0 * h)
    xs = {}
    for side in (0, 2, 1):
        x, st, hist = ev.linear_solve(u, b, sigma, precond=precond, ksp_pc_side=side, ksp_rtol=1e-12,
                                      ksp_restart=200, ksp_max_it=2000)
        print("[pc_side]", precond, side, st["ksp_its"], hist[0], hist[-1], st["fnorm"] / st["fnorm0"], flush=True)
        assert st["ksp_its"] < 2000
        xs[side] = x
        if side == 2:
            # flexible GMRES: the returned solution's true residual is the last estimate
            assert st["fnorm"] == pytest.approx(hist[-1], rel=1e-3, abs=1e-13 * hist[0])
    # all solve (sigma I + J) x = b
    assert float(torch.linalg.norm(xs[0] - xs[1]) / torch.linalg.norm(xs[0])) <= 1e-8
    assert float(torch.linalg.norm(xs[0] - xs[2]) / torch.linalg.norm(xs[0])) <= 1e-8
    assert hist[0] != pytest.approx(float(torch.linalg.norm(b)), rel=1e-3)  # left: ||M^-1 b||


def test_lagged_preconditioner_rebuilds_every_k_iterations():
    """pc_lag k rebuilds the preconditioner every k-th Newton iteration,
    counted across steps, and converges to the same implicit steps."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(8, 4, 2)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    dt0 = 3.0 * h / CF.sound_speed_max(u0, cfg)
    opts = dict(jac_type=0, precond=3, ksp_restart=60, ksp_max_it=600, ksp_rtol=1e-8, snes_rtol=1e-9,
                snes_max_it=20)
    ra = ev.pseudo_run(a := u0.clone(), dt0, 1.5, 4, pc_lag=1, **opts)
    rb = ev.pseudo_run(b := u0.clone(), dt0, 1.5, 4, pc_lag=3, **opts)
    sa, sb = ra["stats"], rb["stats"]
    print("[lag]", sa, sb, flush=True)
    assert sa["pc_builds"] == sa["newton_its"]
    assert sb["pc_builds"] == -(-sb["newton_its"] // 3)  # ceil(its / 3)
    assert float(torch.linalg.norm(a - b) / torch.linalg.norm(a)) <= 1e-7
    assert sa["t_pc_factor"] > 0.0 and sa["t_jv"] > 0.0 and sa["t_orth"] > 0.0 and sa["t_pc_apply"] > 0.0


def test_single_precision_ilu_factors_precondition_the_same_solve():
    """pc_fp32: the ILU(0) factors applied in single precision (double
    accumulation) still precondition flexible GMRES to the same solution; the
    Newton step converges to the same state."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(8, 4, 3)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    u0.mul_(1.0 + 0.01 * torch.sin(0.37 * torch.arange(u0.numel(), device="cuda", dtype=torch.float64)))
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    dt = 3.0 * h / CF.sound_speed_max(u0, cfg)
    out, sts = {}, {}
    for fp32 in (0, 1):
        u = u0.clone()
        sts[fp32] = ev.beuler_step(u, dt, jac_type=0, precond=3, ksp_pc_side=2, pc_fp32=fp32, ksp_restart=60,
                                   ksp_rtol=1e-10, snes_rtol=1e-10, snes_max_it=10)
        out[fp32] = u
    print("[fp32 ilu]", sts, flush=True)
    assert sts[0]["converged"] and sts[1]["converged"]
    assert sts[1]["ksp_its"] <= 2 * sts[0]["ksp_its"] + 5
    assert float(torch.linalg.norm(out[0] - out[1]) / torch.linalg.norm(out[0])) <= 1e-9

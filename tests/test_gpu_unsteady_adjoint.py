"""frsdg_unsteady_adjoint (C ABI) against a step-by-step reference built from
the library's single products: the reference schedule of tests/revolve_ref.py
driving an adjoint step composed of frsdg_residual, frsdg_vjp and
frsdg_filter_transpose with the transposed SSP-RK3 combinations in torch
(the algebra written out in fr_sdg_amd/adjoint.py).  The native sweep must give
the same gradient (1e-13: fused outputs round differently) after exactly the
same number of recomputed forward steps; the FD check of the gradient itself is
test_gpu_solvers.test_unsteady_adjoint_rho2_vs_fd."""
import pytest

import cases as C
import revolve_ref as RR

pytestmark = pytest.mark.gpu


class TorchStepAdjoint(RR.Revolve):
    def adjoint_step(self, u, lam):
        ev, dt = self.ev, self.dt
        u1 = u - dt * ev.residual_kernel(u)
        u2 = 0.75 * u + 0.25 * (u1 - dt * ev.residual_kernel(u1))
        l3 = ev.filter_transpose(lam) if ev.cfg.get("apply_filter", 0) else lam
        l2 = (2.0 / 3.0) * (l3 - dt * ev.vjp(u2, l3))
        l1 = 0.25 * (l2 - dt * ev.vjp(u1, l2))
        self.adjoint_steps += 1
        return (1.0 / 3.0) * l3 + 0.75 * l2 + l1 - dt * ev.vjp(u, l1)


CASES = {
    "warped_vdw_p3_filter": (C.warped_cfg(6, 5, 3, filt=1), 2e-3, [0.3, 0.5, 0.1, -0.1, 1.0, 1.0]),
    "bump_p2_nofilter": (C.bump_cfg(6, 3, 2), 4e-3, None),
    "sod1d_p3_filter": (dict(C.sod_cfg(40, 3, 1), apply_filter=1, filt_cutoff=0, filt_strength=8, filt_alpha=36.0),
                        5e-4, None),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("T,s", [(1, 4), (9, 0), (17, 2), (23, 32)])
def test_native_unsteady_adjoint_matches_reference(name, T, s):
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfg, dt, pulse = CASES[name]
    print(f"[unsteady-adjoint] {name} T={T} s={s}", flush=True)
    ev = fr.PhysicsEvaluator(cfg)
    if cfg["dims"] == 1:
        u0 = ev.impose_ic(fr.IC_SOD, CF.SOD_PRM)
    elif pulse:
        u0 = ev.impose_ic(fr.IC_PULSE, pulse)
    else:
        u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
        u0.mul_(1.0 + 0.01 * torch.sin(0.3 * torch.arange(u0.numel(), device="cuda", dtype=torch.float64)))
    lamT = torch.cos(0.7 * torch.arange(u0.numel(), device="cuda", dtype=torch.float64))
    ref = TorchStepAdjoint(ev, dt, T, max_snapshots=s)
    g_ref = ref.gradient(u0, lamT)
    u0c = u0.clone()
    g, st = ev.unsteady_adjoint(u0, lamT, dt, T, s)
    torch.cuda.synchronize()
    assert torch.equal(u0, u0c)
    assert st["adjoint_steps"] == T == ref.adjoint_steps
    assert st["forward_steps"] == ref.forward_steps
    assert st["snapshots_used"] == ref.peak_snapshots <= s
    rel = float(torch.linalg.norm(g - g_ref) / torch.linalg.norm(g_ref))
    assert rel <= 1e-13, rel
    # in place: lam_0 may alias lam_T
    lam = lamT.clone()
    ev.unsteady_adjoint(u0, lam, dt, T, s, out=lam)
    assert torch.equal(lam, g)


def test_unsteady_adjoint_refuses_the_weno_limiter():
    import torch

    import fr_sdg_amd as fr
    cfg = dict(C.explosion_cfg(6, 5, 3), apply_weno=1)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0])
    with pytest.raises(fr.FrsdgError, match="unsupported"):
        ev.unsteady_adjoint(u0, torch.ones_like(u0), 1e-3, 3, 2)

"""Failure detection of explicit stepping (the analogue of the reference
harness's -ts_error_if_step_fails, tools/check_radial.py:208-210): every stage
output is checked where the stage kernel stores it.  The failures are forced
from the host (a state with one bad node written before stepping); every
kernel still runs to completion."""
import math

import numpy as np
import pytest

import cases as C

pytestmark = pytest.mark.gpu


def _setup(cfg, pulse=True):
    import torch

    import fr_sdg_amd as fr
    ev = fr.PhysicsEvaluator(cfg)
    if cfg["dims"] == 1:
        # a smooth 1D state (the plain-method Sod tube itself loses positivity within a few steps)
        u = ev.impose_ic(fr.IC_UNIFORM, [1.0, 0.2, 0.0, 1.0])
        u.mul_(1.0 + 0.05 * torch.sin(0.3 * torch.arange(u.numel(), device="cuda", dtype=torch.float64)))
    else:
        u = ev.impose_ic(fr.IC_PULSE, [0.2, 0.3 * cfg["lx"], 0.5 * cfg["lx"], 0.5 * cfg["ly"], 1.0, 1.0])
    return torch, fr, ev, u


CFGS = {"vortex_p4": C.vortex_cfg(8, 6, 4), "warped_p7_filter": C.warped_cfg(5, 4, 7),
        "sod1d_p3": C.sod_cfg(40, 3, 1)}


@pytest.mark.parametrize("name", sorted(CFGS))
def test_healthy_run_reports_nothing(name):
    torch, fr, ev, u = _setup(CFGS[name])
    assert ev.check_state(u) == 0
    ev.ssprk_run(u, 1e-4, 20)     # graph chunk + plain steps, no error
    ev.ssprk_step(u, 1e-4)
    assert ev.step_health() == 0
    assert ev.check_state(u) == 0


@pytest.mark.parametrize("name", sorted(CFGS))
@pytest.mark.parametrize("bad", ["negative_density", "nan"])
def test_forced_bad_node_is_flagged_and_stops_the_run(name, bad):
    torch, fr, ev, u = _setup(CFGS[name])
    cfg = CFGS[name]
    h = u.cpu().numpy().copy()
    n = cfg["p"] + 1
    # one node in the middle of the mesh: its density (2D layout node*4, 1D [e][eq][node])
    if cfg["dims"] == 2:
        i = (h.size // 4 // 2) * 4
    else:
        i = (cfg["nx"] // 2) * 3 * n + 1
    h[i] = -0.5 if bad == "negative_density" else math.nan
    v = torch.from_numpy(h).cuda()
    want = fr.HEALTH_DENSITY if bad == "negative_density" else fr.HEALTH_NONFINITE
    assert ev.check_state(v) & want
    # one plain step: the first stage output at that node is bad
    w = v.clone()
    ev.ssprk_step(w, 1e-4)
    assert ev.step_health() != 0
    assert ev.step_health() == 0  # cleared by the previous call
    # the graph-replayed run stops with a numerical failure after the first graph
    # (stop mask: every flag; the default stops on non-finite states only)
    ev.set_stop_mask(fr.HEALTH_NONFINITE | fr.HEALTH_DENSITY | fr.HEALTH_ENERGY)
    x = v.clone()
    with pytest.raises(fr.FrsdgError, match="error 5"):
        ev.ssprk_run(x, 1e-4, 48)


@pytest.mark.parametrize("name", sorted(CFGS))
def test_default_stop_mask_runs_through_non_positive_nodes(name):
    # a finite state with one node of negative internal energy: flagged, but
    # under the default mask (non-finite only) one graph of steps runs through
    # unless the state itself turns non-finite
    torch, fr, ev, u = _setup(CFGS[name])
    cfg = CFGS[name]
    h = u.cpu().numpy().copy()
    n = cfg["p"] + 1
    if cfg["dims"] == 2:
        i = (h.size // 4 // 2) * 4
        h[i + 3] = 0.1 * (h[i + 1] ** 2 + h[i + 2] ** 2) / (2.0 * h[i]) - 1e-3  # rho E below the kinetic part
    else:
        i = (cfg["nx"] // 2) * 3 * n + 2 * n + 1
        h[i] = -1e-3
    v = torch.from_numpy(h).cuda()
    assert ev.check_state(v) & fr.HEALTH_ENERGY
    x = v.clone()
    try:
        ev.ssprk_run(x, 1e-6, 16)
        nonfinite = False
    except fr.FrsdgError:
        nonfinite = True
    assert bool(ev.check_state(x) & fr.HEALTH_NONFINITE) == nonfinite
    with pytest.raises(fr.FrsdgError, match="invalid"):
        ev.set_stop_mask(8)


@pytest.mark.parametrize("bad", ["negative_density", "nan"])
def test_batched_sweep_flags_the_instance_with_a_bad_node(bad):
    """Batched sweep: per-instance flag words; a bad node forced from the host
    into one instance is flagged for that instance alone, and a batched run
    stops after the graph in which an instance raised a flag of its stop mask
    (non-finite by default; here the density bit too for that instance)."""
    import torch

    import fr_sdg_amd as fr
    from fr_sdg_amd import configs as CF
    cfgs = [CF.explosion(6, 2, vdw=False), CF.explosion(5, 4, vdw=True), CF.explosion(4, 6, vdw=False)]
    b = fr.Batch(cfgs)
    u = b.empty()
    # the radial jump spread over a tanh layer wider than these coarse elements
    for k, iv in enumerate(b.instances):
        iv.impose_ic(fr.IC_EXPLOSION, CF.EXPLOSION_PRM + [0.6], out=b.view(u, k))
    b.set_dt([1e-4] * len(cfgs))
    good = u.clone()
    b.ssprk_run(good, 10)                  # one graph of steps and plain steps
    assert b.health() == [0, 0, 0]
    # a bad density node in the middle of instance 1
    h = u.cpu().numpy().copy()
    mid = b.offsets[1] + ((b.offsets[2] - b.offsets[1]) // 8) * 4
    h[mid] = -0.5 if bad == "negative_density" else math.nan
    v = torch.from_numpy(h).cuda()
    w = v.clone()
    b.ssprk_step(w)
    f = b.health(clear=True)
    want = fr.HEALTH_DENSITY if bad == "negative_density" else fr.HEALTH_NONFINITE
    assert f[1] & want and f[0] == 0 and f[2] == 0, f
    assert b.health() == [0, 0, 0]
    # the run stops on the instance's stop mask
    b.instances[1].set_stop_mask(fr.HEALTH_NONFINITE | fr.HEALTH_DENSITY)
    w = v.clone()
    with pytest.raises(fr.FrsdgError):
        b.ssprk_run(w, 16)
    assert b.health()[1] & want
    b.instances[1].set_stop_mask(fr.HEALTH_NONFINITE)
    # the other instances step as in the healthy run (instances do not interact)
    g = u.clone()
    b.ssprk_step(g)
    w = v.clone()
    b.ssprk_step(w)
    for k in (0, 2):
        assert torch.equal(b.view(w, k), b.view(g, k)), k

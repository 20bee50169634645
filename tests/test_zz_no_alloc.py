"""The allocation-free compute_error entry point (frsdg_l2_error_noalloc),
called through the C ABI, against frsdg_l2_error on the same inputs: two calls
in a row on one GPU, the second on a smaller problem (the static scratch keeps
the first call's partial sums), with and without the log-variable state, and
the size limit."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import fr_sdg_amd as fr  # noqa: E402
from fr_sdg_amd import configs as CF  # noqa: E402


def _call(lib, name, ev, kind, prm, u, t):
    p = (ctypes.c_double * 8)(*(list(prm) + [0.0] * (8 - len(prm))))
    out = (ctypes.c_double * 4)()
    rc = getattr(lib, name)(ev.h, kind, p, t, ctypes.c_void_p(u.data_ptr()), out,
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, list(out)


def _perturbed(ev, kind, prm, seed):
    u = ev.impose_ic(kind, prm)
    g = torch.Generator(device="cpu").manual_seed(seed)
    return u + 1e-4 * torch.randn(u.numel(), generator=g, dtype=torch.float64).to(u.device)


CASES = [
    ("warped_vdw_p7", dict(CF.warped(48), apply_filter=0), fr.IC_EXPLOSION, CF.WARPED_PRM, 0.0),
    ("vortex_p4", CF.vortex(32, 24), fr.IC_VORTEX, CF.VORTEX_PRM, 0.3),
    ("vortex_logvars_p3", dict(CF.vortex(20, 12, 3), cond_sys=1), fr.IC_VORTEX, CF.VORTEX_PRM, 0.0),
    ("bump_p5_small", CF.bump(16, 6, 5), fr.IC_UNIFORM, CF.bump_freestream(), 0.0),
]


@pytest.mark.parametrize("first,second", [(0, 1), (1, 2), (2, 3), (0, 3)])
def test_noalloc_l2_error_matches_allocating_entry_point(first, second):
    lib = fr.load_library()
    # the second problem is smaller than the first in every pair
    evs = []
    for idx in (first, second):
        name, cfg, kind, prm, t = CASES[idx]
        ev = fr.PhysicsEvaluator(cfg)
        u = _perturbed(ev, kind, prm, seed=idx + 11)
        evs.append((name, ev, kind, prm, u, t))
    assert evs[1][1].ndofs < evs[0][1].ndofs
    for name, ev, kind, prm, u, t in evs:
        print(f"[no-alloc] {name} ndofs={ev.ndofs}", flush=True)
        rc_a, ref = _call(lib, "frsdg_l2_error", ev, kind, prm, u, t)
        rc_b, got = _call(lib, "frsdg_l2_error_noalloc", ev, kind, prm, u, t)
        assert rc_a == 0 and rc_b == 0
        assert np.all(np.isfinite(ref)) and min(ref) > 0.0
        # same kernels, same order of summation: bitwise equal
        assert got == ref, (name, got, ref)


def test_noalloc_l2_error_repeated_calls_do_not_accumulate():
    lib = fr.load_library()
    name, cfg, kind, prm, t = CASES[1]
    ev = fr.PhysicsEvaluator(cfg)
    u = _perturbed(ev, kind, prm, seed=3)
    r1 = _call(lib, "frsdg_l2_error_noalloc", ev, kind, prm, u, t)
    r2 = _call(lib, "frsdg_l2_error_noalloc", ev, kind, prm, u, t)
    assert r1 == r2 and r1[0] == 0


def test_noalloc_l2_error_rejects_more_nodes_than_its_scratch():
    lib = fr.load_library()
    # 1024 x 1025 elements of 16 nodes = 16,793,600 nodes > 2^24: rejected before any kernel
    ev = fr.PhysicsEvaluator(dict(dims=2, nx=1024, ny=1025, p=3, lx=1.0, ly=1.0, bc_type=[1, 1, 1, 1]))
    u = torch.zeros(8, dtype=torch.float64, device="cuda")
    rc, _ = _call(lib, "frsdg_l2_error_noalloc", ev, fr.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0], u, 0.0)
    assert rc == 1
    # exactly 2^24 nodes is within the limit
    ev2 = fr.PhysicsEvaluator(dict(dims=2, nx=1024, ny=1024, p=3, lx=1.0, ly=1.0, bc_type=[1, 1, 1, 1]))
    u2 = ev2.impose_ic(fr.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0])
    rc_a, ref = _call(lib, "frsdg_l2_error", ev2, fr.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0], u2, 0.0)
    rc_b, got = _call(lib, "frsdg_l2_error_noalloc", ev2, fr.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0], u2, 0.0)
    assert rc_a == 0 and rc_b == 0 and got == ref

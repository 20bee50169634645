"""Benchmark: FP64 DOF-RHS evals/s of the FR residual on one MI355X.

Headline workload (one "step" = one SSP-RK3 time step = 3 launches of the
fused residual + stage kernel k_res2d_stage<8, curved>, the post-step
spectral filter fused into the third): the large curved radial shock tube,
[-1,1]^2 warped by x += 0.05 sin(pi x) sin(pi y) (same in y), 1024^2 quads,
p=7 (64 nodes/element, 268,435,456 DOFs, 2.15 GB per state), van der Waals
gas (a=0.1, b=0.02), HLLC, exponential filter (order 8, alpha 36), CFL
0.4/(2p+1).  The stepping graph (16 steps) is captured, instantiated and
replayed during the warm-up; the timed region replays it (plus plain
launches for steps % 16), inputs resident in HBM.

    python bench.py [--gpus 1] [--steps 200] [--warmup 32]

Prints ONE JSON line.  Extra fields: the 2D isentropic vortex (128^2, p=4,
HLLC, periodic) stepped the same way; the residual-only and J*v kernels;
the JFNK comparison on the bump; the batched sweep; measured peaks (HBM copy
of a 2.15 GB state, FP64 FMA loop); the CPU baseline (the oracle on the same
state, all host threads and one thread).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time


def progress(msg):
    """One line per finished leg on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 DOF-RHS evals/s + Newton step time (tangent vs complex-step), 1 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
GRAPH_CHUNK = 16  # steps per captured graph in frsdg_ssprk_run


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-large", action="store_true", help="skip the residual-only / limiter figures")
    ap.add_argument("--no-newton", action="store_true", help="skip the implicit (JFNK) comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep comparison")
    ap.add_argument("--headline-only", action="store_true", help="only the headline step (profiling runs)")
    ap.add_argument("--newton-configured", action="store_true",
                    help="add the configured bump schedule from the free stream (CFL 10, x1.1, Newton rtol 1e-8), "
                         "two steps per product (about a minute)")
    ap.add_argument("--ilu", action="store_true",
                    help="add the coloured-matrix / ILU(0) preconditioner timings on the bump (about 30 s)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-core CPU baseline (0: OMP_NUM_THREADS, else min(16, cores))")
    ap.add_argument("--adjoint-large", action="store_true",
                    help="run only the unsteady adjoint of int(rho^2) on the 1024^2 p=7 warped mesh")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--adj-snapshots", type=int, default=32)
    ap.add_argument("--adj-n", type=int, default=1024)
    ap.add_argument("--adj-fd-eps", default="1e-2,1e-3,1e-4",
                    help="comma-separated steps of the central finite difference that checks the adjoint")
    ap.add_argument("--kernel-only", choices=["step", "stage", "residual", "tangent", "cstep", "adjoint", "large",
                                                "tangent_large", "adjoint_large",
                                                "limiter", "copy", "vcopy", "batch", "batch_each", "ilu"],
                    default=None,
                    help="profiling mode: launch only this kernel `--steps` times, print nothing else")
    return ap.parse_args()


def vortex_cfg(nx=128, ny=128, p=4):
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=10.0, ly=10.0, riemann=1, bc_type=[5, 5, 5, 5])


VORTEX_PRM = [1.0, 1.0, 5.0, 1.0, 10.0]


def flops_per_element(n):
    # residual, HLLC / ideal gas: 32 n^3 + 89 n^2 + 424 n
    return 32 * n ** 3 + 89 * n ** 2 + 424 * n


def time_kernel(torch, fn, reps):
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / 1e3 / reps


def stepping_setup(torch, fr, which):
    """(evaluator, initial state, dt, cfg, workload text) of a stepped workload:
    'large' = the curved radial shock tube (headline), 'vortex' = the periodic
    isentropic vortex.  dt = CFL * h / max(|v| + c) at t = 0, CFL 0.4/(2p+1)."""
    from fr_sdg_amd import configs as CF
    if which == "large":
        cfg = CF.warped(1024)
        ev = fr.PhysicsEvaluator(cfg)
        u0 = ev.impose_ic(fr.IC_EXPLOSION, CF.WARPED_PRM)
        h = cfg["lx"] / cfg["nx"]
        text = ("curved radial shock tube: [-1,1]^2 warped by 0.05 sin(pi x) sin(pi y), 1024x1024 quads, p=7, "
                "van der Waals a=0.1 b=0.02, HLLC, exponential filter order 8 alpha 36 (post-step, fused), "
                "SSP-RK3 at CFL 0.4/(2p+1); jump r=0.4 (1,1)/(0.125,0.1) spread over a tanh layer of width 0.1")
    else:
        cfg = vortex_cfg()
        ev = fr.PhysicsEvaluator(cfg)
        u0 = ev.impose_ic(fr.IC_VORTEX, VORTEX_PRM)
        h = cfg["lx"] / cfg["nx"]
        text = "2D isentropic vortex 128x128 p=4 HLLC ideal, periodic, SSP-RK3 at CFL 0.4/(2p+1)"
    dt = CF.explicit_cfl(cfg["p"]) * h / CF.sound_speed_max(u0, cfg)
    return ev, u0, dt, cfg, text


def timed_steps(torch, ev, u, dt, steps, warmup):
    """Warm up with whole graph chunks (the graph of these exact u/dt is
    captured and instantiated there), then time `steps` steps with HIP events
    on the launch stream.  Returns (seconds, wall seconds)."""
    w = max(GRAPH_CHUNK, GRAPH_CHUNK * ((warmup + GRAPH_CHUNK - 1) // GRAPH_CHUNK))
    ev.ssprk_run(u, dt, w)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall0 = time.perf_counter()
    e0.record(s)
    ev.ssprk_run(u, dt, steps)
    e1.record(s)
    e1.synchronize()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3, time.perf_counter() - t_wall0


def step_bytes_per_dof(n, curved):
    """Algorithmic HBM bytes of one SSP-RK3 step per DOF: stage 1 reads u and
    writes w1 (16), stages 2 and 3 also read u^n (24 each); curved meshes add
    per launch the node metrics (8) and the face metrics (nx, ny, scale per
    face node: 12/n)."""
    return 64.0 + (3.0 * (8.0 + 12.0 / n) if curved else 0.0)


def load_peaks_lib():
    from fr_sdg_amd import build as B
    if not os.path.exists(B.PEAKS_LIB):
        raise RuntimeError(f"{B.PEAKS_LIB} not built; run __graft_entry__.build()")
    L = ctypes.CDLL(B.PEAKS_LIB)
    L.frsdg_peak_fma.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.frsdg_peak_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int,
                                  ctypes.c_void_p]
    return L


def measured_peaks(torch):
    """HBM copy bandwidth on a 2.15 GB state (the headline's state size; read +
    write counted) and the FP64 vector FMA-loop rate, beside the datasheet
    figures."""
    L = load_peaks_lib()
    s = torch.cuda.current_stream()
    st = ctypes.c_void_p(s.cuda_stream)
    n = 1024 * 1024 * 64 * 4  # doubles of one headline state
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    blocks = 256 * 8 * 4
    run = lambda: L.frsdg_peak_copy(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), n, blocks, st)
    if run():
        raise RuntimeError("copy probe failed")
    t_copy = time_kernel(torch, run, 20)
    t_torch = time_kernel(torch, lambda: b.copy_(a), 20)
    del a, b
    fblocks, iters = 256 * 4 * 8, 4096
    o = torch.empty(fblocks * 256, dtype=torch.float64, device="cuda")
    runf = lambda: L.frsdg_peak_fma(ctypes.c_void_p(o.data_ptr()), fblocks, iters, st)
    if runf():
        raise RuntimeError("fma probe failed")
    t_fma = time_kernel(torch, runf, 10)
    fl = 16.0 * iters * fblocks * 256
    del o
    torch.cuda.empty_cache()
    return {"copy_GBps": 16.0 * n / t_copy / 1e9, "copy_torch_GBps": 16.0 * n / t_torch / 1e9,
            "copy_bytes": 16.0 * n, "fp64_fma_TFLOPs": fl / t_fma / 1e12,
            "datasheet": {"hbm_GBps": HBM_PEAK_GBS, "fp64_vector_TFLOPs": FP64_PEAK_TFS}}


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (JSON)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch"), name


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "libfrsdg runs on one GPU only (no multi-GPU path); use --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import fr_sdg_amd as fr

    torch.cuda.set_device(0)
    if args.kernel_only:
        return kernel_only(torch, fr, args)
    if args.adjoint_large:
        print(json.dumps(adjoint_large(torch, fr, args)))
        return

    # ---------------- headline: curved radial shock tube, graph-replayed SSP-RK3 ----------------
    ev, u0, dt, cfg, text = stepping_setup(torch, fr, "large")
    n = cfg["p"] + 1
    E = cfg["nx"] * cfg["ny"]
    ndof = ev.ndofs
    u = u0.clone()
    elapsed, t_wall = timed_steps(torch, ev, u, dt, args.steps, args.warmup)
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError("non-finite state after the timed steps")
    n_res = 3 * args.steps
    value = ndof * n_res / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    bytes_launch = step_bytes_per_dof(n, True) * ndof / 3.0
    t_launch = elapsed / n_res
    achieved = bytes_launch / t_launch / 1e9
    # FP64 work per launch: the residual (HLLC, ideal-gas count) + the filter's
    # two line sweeps (16 n^3 per element) once per step
    flops_launch = (flops_per_element(n) + 16 * n ** 3 / 3.0) * E
    traffic, tsrc = read_traffic("traffic_step_warped_p7_1024.json")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": text, "ndofs": ndof, "elements": E, "p": cfg["p"], "dt": dt,
                   "residual_evals_per_step": 3, "graph_captured": True,
                   "graph_steps": GRAPH_CHUNK, "warmup_steps_run": max(GRAPH_CHUNK, GRAPH_CHUNK * (
                       (args.warmup + GRAPH_CHUNK - 1) // GRAPH_CHUNK))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_res2d_stage<8,curved> (residual + SSP-RK stage; the filter fused into every "
                               "third launch), 3 launches per step",
                     "bytes_per_launch_alg": bytes_launch, "bytes_per_dof_step": step_bytes_per_dof(n, True),
                     "avg_launch_s": t_launch, "traffic_source": tsrc,
                     "fp64_tflops": flops_launch / t_launch / 1e12,
                     "fp64_frac": flops_launch / t_launch / 1e12 / FP64_PEAK_TFS},
        "wall_s_timed": t_wall,
    }
    del u
    progress(f"headline done: {ms_per_step:.3f} ms/step, {achieved / HBM_PEAK_GBS:.3f} of HBM peak")
    if not args.headline_only:
        extra_legs(torch, fr, args, line, ev, u0)
    del ev, u0
    torch.cuda.empty_cache()
    if not args.no_cpu and not args.headline_only:
        line["cpu_baseline"] = cpu_baseline(args, dt)
        progress("cpu baseline done")
    print(json.dumps(line))


def extra_legs(torch, fr, args, line, ev, u0):
    """Extra keys of the bench line (each leg reports its error instead of failing the line)."""
    def leg(key, fn):
        try:
            line[key] = fn()
        except Exception as exc:  # report, keep the headline line valid
            line[key] = {"error": repr(exc)}
        progress(f"{key} done")

    leg("measured_peaks", lambda: measured_peaks(torch))
    if not args.no_large:
        leg("large_curved", lambda: bench_large(torch, fr, ev, u0))
    leg("vortex", lambda: bench_vortex(torch, fr, args))
    if not args.no_newton:
        leg("newton", lambda: bench_newton(torch, fr))
        if args.ilu:
            leg("ilu", lambda: bench_ilu(torch, fr))
        if args.newton_configured:
            leg("newton_configured", lambda: bench_newton_configured(torch, fr))
    leg("widened_rows", lambda: bench_rows(torch, fr))
    if not args.no_batch:
        leg("batched", lambda: bench_batch(torch, fr))
        # the same sweep on meshes 1/4 the size (N in 8..48): launch-bound instances
        # (20 steps: within 100 steps shocks form and the filtered p=5,6
        # instances lose positivity on these coarse meshes)
        leg("batched_small", lambda: bench_batch(torch, fr, steps=20, scale=0.25))


def bench_vortex(torch, fr, args):
    """The periodic isentropic vortex (128^2, p=4, affine, 13 MB per state:
    cache-resident) stepped like the headline, plus its residual-only and
    tangent / complex-step J*v kernels."""
    ev, u0, dt, cfg, text = stepping_setup(torch, fr, "vortex")
    ndof = ev.ndofs
    u = u0.clone()
    steps = max(args.steps, 160)
    elapsed, _ = timed_steps(torch, ev, u, dt, steps, args.warmup)
    t_launch = elapsed / (3 * steps)
    v = torch.randn(ndof, dtype=torch.float64, device="cuda")
    out = ev.empty()
    t_tan = time_kernel(torch, lambda: ev.jvp(u, v, 1.0, out=out), 50)
    t_cs = time_kernel(torch, lambda: ev.jvp_cstep(u, v, 1.0, 1e-7, out=out), 50)
    t_vjp = time_kernel(torch, lambda: ev.vjp(u, v, 1.0, out=out), 50)
    t_res = time_kernel(torch, lambda: ev.residual_kernel(u, out=out), 100)
    bytes_launch = step_bytes_per_dof(5, False) * ndof / 3.0
    res = {"workload": text, "ndofs": ndof, "steps": steps, "dt": dt, "ms_per_step": 1e3 * elapsed / steps,
           "dof_rhs_per_s": 3 * ndof * steps / elapsed, "avg_launch_s": t_launch,
           "hbm_frac_alg": bytes_launch / t_launch / 1e9 / HBM_PEAK_GBS,
           "residual_only": {"avg_s": t_res, "dof_rhs_per_s": ndof / t_res, "gbs_alg": 16.0 * ndof / t_res / 1e9},
           "jvp": {"tangent_s": t_tan, "complex_step_s": t_cs, "speedup": t_cs / t_tan, "adjoint_s": t_vjp},
           "finite": bool(torch.isfinite(u).all())}
    del ev, u, u0, v, out
    torch.cuda.empty_cache()
    return res


def kernel_only(torch, fr, args):
    """Launch one kernel kind back to back (for rocprofv3 counter / trace runs).
    'step' = SSP-RK3 steps of the headline (plain launches, no graph), 'copy' =
    a plain copy of the headline's 2.15 GB state, 'vcopy' = of the vortex's
    13 MB state (counter scaling references)."""
    from fr_sdg_amd import configs as CF
    if args.kernel_only in ("batch", "batch_each"):
        b, u = batch_setup(fr)
        b.set_tuning(0, 1, 1)  # one workgroup per tile, one launch per order
        step = b.ssprk_step if args.kernel_only == "batch" else b.ssprk_step_each
        for _ in range(args.steps):
            step(u)
        torch.cuda.synchronize()
        print(json.dumps({"kernel_only": args.kernel_only, "steps": args.steps, "ndofs": b.ndofs}))
        return
    if args.kernel_only == "ilu":
        # the bump 256x64 p=5: one coloured assembly (FD), one ILU(0) factorisation,
        # then `steps` applies
        cfg = CF.bump(tangent_bc_mode=0)
        ev = fr.PhysicsEvaluator(cfg)
        u = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
        sigma = CF.sound_speed_max(u, cfg) / (10.0 * min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"]))
        P = ev.color_jacobian(u, sigma, None, method=0)
        Dinv = ev.ilu0_factor(P)
        x = torch.randn(ev.ndofs, dtype=torch.float64, device="cuda")
        for _ in range(args.steps):
            ev.ilu0_apply(P, Dinv, x)
        torch.cuda.synchronize()
        print(json.dumps({"kernel_only": "ilu", "applies": args.steps, "ndofs": ev.ndofs}))
        return
    k = args.kernel_only
    if k in ("step", "large", "limiter", "copy", "tangent_large", "adjoint_large"):
        ev, u, dt, cfg, _ = stepping_setup(torch, fr, "large")
        if k == "limiter":
            ev = fr.PhysicsEvaluator(dict(cfg, apply_filter=0, apply_weno=1))
    else:
        ev, u, dt, cfg, _ = stepping_setup(torch, fr, "vortex")
    out = ev.empty()
    v = torch.randn(ev.ndofs, dtype=torch.float64, device="cuda")
    fn = {
        "step": lambda: ev.ssprk_step(u, dt),
        "stage": lambda: ev.ssprk_step(u, dt),
        "residual": lambda: ev.residual_kernel(u, out=out),
        "large": lambda: ev.residual_kernel(u, out=out),
        "limiter": lambda: ev.limit_solution(u, out=out),
        "tangent": lambda: ev.jvp(u, v, 1.0, out=out),
        "cstep": lambda: ev.jvp_cstep(u, v, 1.0, 1e-7, out=out),
        "adjoint": lambda: ev.vjp(u, v, 1.0, out=out),
        "tangent_large": lambda: ev.jvp(u, v, 1.0, out=out),
        "adjoint_large": lambda: ev.vjp(u, v, 1.0, out=out),
        "copy": lambda: out.copy_(u),
        "vcopy": lambda: out.copy_(u),
    }[k]
    for _ in range(args.steps):
        fn()
    torch.cuda.synchronize()
    print(json.dumps({"kernel_only": k, "launches": args.steps, "ndofs": ev.ndofs,
                      "bytes_per_copy": 16 * ev.ndofs if k in ("copy", "vcopy") else None}))


def bench_large(torch, fr, ev, u):
    """Residual only (no stage, no filter) and the WENO limiter on the
    headline's initial state (1024^2 warped p=7 VdW HLLC, 2.15 GB per state)."""
    R = ev.empty()
    t = time_kernel(torch, lambda: ev.residual_kernel(u, out=R), 20)
    n = 8
    ndof = ev.ndofs
    E = ev.cfg["nx"] * ev.cfg["ny"]
    bytes_alg = (16.0 + 8.0 + 12.0 / n) * ndof  # u, R, volume metrics, face (nx,ny,scale)
    ok = bool(torch.isfinite(R).all())
    res = {"workload": "warped 1024x1024 p=7 VdW HLLC residual", "ndofs": ndof, "avg_s": t,
           "dof_rhs_per_s": ndof / t, "gbs_alg": bytes_alg / t / 1e9,
           "hbm_frac": bytes_alg / t / 1e9 / HBM_PEAK_GBS,
           "fp64_tflops": flops_per_element(n) * E / t / 1e12, "finite": ok}
    # derivative products on the same state: forward tangent (24 B/DOF + metrics),
    # complex step, adjoint (u, w, out, lifted face cotangents: k_adj_lift + k_adj2d_tile)
    v = torch.randn(ndof, dtype=torch.float64, device="cuda")
    t_tan = time_kernel(torch, lambda: ev.jvp(u, v, 1.0, out=R), 10)
    t_cs = time_kernel(torch, lambda: ev.jvp_cstep(u, v, 1.0, 1e-7, out=R), 10)
    t_vjp = time_kernel(torch, lambda: ev.vjp(u, v, 1.0, out=R), 10)
    res["derivatives"] = {"tangent_s": t_tan, "complex_step_s": t_cs, "adjoint_s": t_vjp,
                          "tangent_over_residual": t_tan / t, "complex_step_over_residual": t_cs / t,
                          "adjoint_over_residual": t_vjp / t,
                          "tangent_gbs_alg": (bytes_alg + 8.0 * ndof) / t_tan / 1e9}
    del R, v
    torch.cuda.empty_cache()
    # WENO limiter with shock sensor on the same state (limit_solution with
    # -apply_weno, no filter): smoothness of every element line, then limiting.
    # Algorithmic bytes: read u, write the output (16 B/DOF); arithmetic:
    # 2 (n-1)(2n+3) unfused FP64 operations per DOF for the smoothness sums.
    evl = fr.PhysicsEvaluator(dict(ev.cfg, apply_filter=0, apply_weno=1))
    out = evl.empty()
    tl = time_kernel(torch, lambda: evl.limit_solution(u, out=out), 20)
    changed = float((out != u).double().mean())
    res["limiter"] = {"workload": "WENO limiter + shock sensor (gamma 0.998, theta 1, threshold 0.8) on the "
                                  "initial warped-mesh explosion state", "avg_s": tl,
                      "dof_per_s": ndof / tl, "gbs_alg": 16.0 * ndof / tl / 1e9,
                      "hbm_frac": 16.0 * ndof / tl / 1e9 / HBM_PEAK_GBS,
                      "fp64_tops_unfused": 2 * (n - 1) * (2 * n + 3) * ndof / tl / 1e12,
                      "fraction_of_dofs_limited": changed}
    del evl, out
    torch.cuda.empty_cache()
    return res


def ilu_flops(nx, ny, nd):
    """FP64 flops of one block ILU(0) factorisation on a non-periodic nx x ny
    box-stencil block pattern: per block row, one product with U_KK^-1 per
    lower neighbour K, one update product per (K, J) pair, one product with
    L_II^-1 per upper neighbour (2 nd^3 each), the unpivoted LU (2/3 nd^3) and
    the two triangular inverses (2/3 nd^3)."""
    def nb(i, j):
        return [(i + di, j + dj) for dj in (-1, 0, 1) for di in (-1, 0, 1)
                if 0 <= i + di < nx and 0 <= j + dj < ny]
    idx = lambda e: e[1] * nx + e[0]
    gemms = 0
    for j in range(ny):
        for i in range(nx):
            me = idx((i, j))
            pat = {idx(e) for e in nb(i, j)}
            for K in (e for e in nb(i, j) if idx(e) < me):
                gemms += 1 + len([J for J in pat if J > idx(K) and J in {idx(f) for f in nb(*K)}])
            gemms += len([J for J in pat if J > me])
    return gemms * 2.0 * nd ** 3 + nx * ny * (4.0 / 3.0) * nd ** 3


def bench_ilu(torch, fr):
    """The reference's preconditioning matrix and its ILU(0) on the bump
    256x64 p=5 (frsdg_color_jacobian: DMDA box colouring, 9 x 144 coloured
    residual evaluations; frsdg_ilu0_factor / frsdg_ilu0_apply): wall times,
    the factorisation's FP64 rate and the apply's HBM rate (it reads every
    stored block once: E*9 blocks of P plus 2 E inverse blocks), and the
    GMRES iterations one CFL-10 backward-Euler Newton iteration from free
    stream needs to reach 1e-3 with each preconditioner (GMRES(60), at most
    3000 iterations)."""
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(tangent_bc_mode=0)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    sigma = CF.sound_speed_max(u0, cfg) / (10.0 * h)
    E, S, nd = ev.color_layout()

    def wall(f, n=1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            r = f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n, r

    out = {"workload": "bump 256x64 p=5 HLLC (2.36 M DOFs), P at free stream, sigma = 1/dt at CFL 10",
           "blocks": E * S, "block": [nd, nd], "P_GB": E * S * nd * nd * 8 / 1e9}
    t_fd, _ = wall(lambda: ev.color_jacobian(u0, sigma, None, method=0))
    t_ex, P = wall(lambda: ev.color_jacobian(u0, sigma, None, method=2))
    t_f, Dinv = wall(lambda: ev.ilu0_factor(P))
    x = torch.randn(ev.ndofs, dtype=torch.float64, device="cuda")
    ev.ilu0_apply(P, Dinv, x)
    t_a, _ = wall(lambda: ev.ilu0_apply(P, Dinv, x), 5)
    t_at, _ = wall(lambda: ev.ilu0_apply_transpose(P, Dinv, x), 3)
    fl = ilu_flops(cfg["nx"], cfg["ny"], nd)
    by = (E * S + 2 * E) * nd * nd * 8.0 + 4.0 * ev.ndofs * 8.0
    out.update({"assemble_fd_s": t_fd, "assemble_exact_s": t_ex, "factor_s": t_f, "factor_gflop": fl / 1e9,
                "factor_tflops": fl / t_f / 1e12, "factor_frac_fp64_peak": fl / t_f / 78.6e12,
                "apply_ms": t_a * 1e3, "apply_GBps": by / t_a / 1e9, "apply_frac_hbm": by / t_a / 8.0e12,
                "apply_transpose_ms": t_at * 1e3,
                "levels": cfg["nx"] + 2 * cfg["ny"] - 2})
    del P, Dinv
    torch.cuda.empty_cache()
    its = {}
    # the last two: flexible GMRES with the ILU(0) factors applied in double and
    # in single precision (pc_fp32: half the bytes per application)
    for name, pc, extra in (("block_jacobi", 1, {}), ("gauss_seidel", 2, {}), ("ilu_fd", 3, {}), ("ilu_exact", 4, {}),
                            ("ilu_fd_fgmres", 3, {"ksp_pc_side": 2}),
                            ("ilu_fd_fgmres_fp32", 3, {"ksp_pc_side": 2, "pc_fp32": 1})):
        u = u0.clone()
        t, st = wall(lambda: ev.beuler_step(u, 1.0 / sigma, jac_type=0, precond=pc, ksp_restart=60,
                                            ksp_max_it=3000, ksp_rtol=1e-3, snes_rtol=1e-3, snes_max_it=1, **extra))
        its[name] = {"gmres_its": st["ksp_its"], "s": t,
                     "apply_ms_per_it": 1e3 * st["t_pc_apply"] / max(1, st["ksp_its"])}
        progress(f"ilu: one Newton iteration with {name}: {st['ksp_its']} GMRES iterations")
    out["one_newton_iteration"] = its
    del ev, u0
    torch.cuda.empty_cache()
    return out


def newton_schedule(torch, fr, ev, u0, jt, start_cfl, inc, cap_cfl, steps, wall, opts, label):
    """Pseudo-transient continuation on the bump driven one step at a time
    (frsdg_pseudo_run with one step of the size the SER rule gives:
    dt_k = inc * dt_0 * f_0 / f_k, -ts_pseudo_increment_dt_from_initial_dt,
    capped at cap_cfl), so every step prints a progress line.  Returns the
    steady-residual history and the summed Newton statistics."""
    from fr_sdg_amd import configs as CF
    cfg = ev.cfg
    h = min(cfg["lx"] / cfg["nx"], cfg["ly"] / cfg["ny"])
    cfl_dt = h / CF.sound_speed_max(u0, cfg)
    u = u0.clone()
    f0 = float(torch.linalg.norm(ev.fun_implicit(u, torch.zeros_like(u))))
    f, dt0 = f0, start_cfl * cfl_dt
    hist = {"reduction": [1.0], "seconds": [0.0], "cfl": [], "newton_its": [], "ksp_its": []}
    tot = {}
    t0 = time.perf_counter()
    for k in range(steps):
        dt = min(inc * dt0 * f0 / f, cap_cfl * cfl_dt)
        r = ev.pseudo_run(u, dt, 1.0, 1, jac_type=jt, **opts)
        f = r["fnorm"][0]
        st = r["stats"]
        for key in ("newton_its", "ksp_its", "pc_builds", "jv_evals", "t_pc_assemble", "t_pc_factor", "t_pc_apply",
                    "t_jv", "t_orth", "seconds"):
            tot[key] = tot.get(key, 0) + st[key]
        el = time.perf_counter() - t0
        hist["reduction"].append(f / f0)
        hist["seconds"].append(el)
        hist["cfl"].append(dt / cfl_dt)
        hist["newton_its"].append(st["newton_its"])
        hist["ksp_its"].append(st["ksp_its"])
        progress(f"newton {label}: step {k}, CFL {dt / cfl_dt:.3g}, steady residual x{f / f0:.3e}, "
                 f"{st['newton_its']} Newton / {st['ksp_its']} GMRES iterations, {el:.1f} s")
        if el > wall or not math.isfinite(f):
            break
    # per attempted Newton iteration (one preconditioner build each: a failed line
    # search still paid for its linear solve)
    its = max(1, tot.get("pc_builds", 0))
    per_it = {k: tot[k] / its for k in ("t_pc_assemble", "t_pc_factor", "t_pc_apply", "t_jv", "t_orth", "seconds")}
    per_it["ksp_its"] = tot.get("ksp_its", 0) / its
    per_it["other"] = per_it["seconds"] - sum(per_it[k] for k in ("t_pc_assemble", "t_pc_factor", "t_pc_apply",
                                                                   "t_jv", "t_orth"))
    kk = max(1, tot.get("ksp_its", 0))
    per_it["per_gmres_iteration"] = {k: tot[k] / kk for k in ("t_pc_apply", "t_jv", "t_orth")}
    return {**hist, "f0": f0, "totals": tot, "per_newton_iteration": per_it,
            "finite": bool(torch.isfinite(u).all())}


BUMP_NEWTON_OPTS = dict(precond=3, ksp_pc_side=2, ksp_restart=60, ksp_rtol=1e-3, snes_rtol=1e-3, snes_max_it=3,
                        snes_stol=1e-8)


def bench_newton(torch, fr, steps=8, wall=30.0):
    """Tangent vs complex-step JFNK on the channel bump 256x64 p=5 (2.36 M
    DOFs), HLLC, driven toward steady state by pseudo-transient continuation:
    SER rule from CFL 1 (x1.5 from the initial dt), capped at CFL 1000 (the
    configured ramp's end); each backward-Euler step solved by Newton (PETSc's
    bt line search, rtol 1e-3, <= 3 iterations) with flexible GMRES(60) to 1e-3
    preconditioned by ILU(0) of the coloured finite-difference matrix (the
    reference's preconditioning matrix and PC).  The tangent kernel linearises
    the wall and Dirichlet states exactly; the complex step sees them frozen,
    as the reference's lambda BCs are.  Reported: the steady-residual history
    of both, the time each needs to reach the reduction the complex step ends
    with, and where the time of a Newton iteration goes (preconditioner
    assembly, factorisation and applications, J*v products, Gram-Schmidt).
    The configured start (CFL 10 from the free stream) is run by
    --newton-configured: no Newton iteration makes progress there."""
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(tangent_bc_mode=0)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    runs = {}
    for name, jt, kmax in (("tangent", 0, 1000), ("complex_step", 1, 300)):
        runs[name] = newton_schedule(torch, fr, ev, u0, jt, 1.0, 1.5, 1000.0, steps, wall,
                                     dict(BUMP_NEWTON_OPTS, ksp_max_it=kmax), name)
    target = min(runs["complex_step"]["reduction"])

    def t_to(run, tgt):
        for red, t in zip(run["reduction"], run["seconds"]):
            if red <= tgt:
                return t
        return None

    tt, tc = t_to(runs["tangent"], target), t_to(runs["complex_step"], target)
    out = {"workload": "bump 256x64 p=5 HLLC (2.36 M DOFs), pseudo-transient continuation (TSPSEUDO SER rule from CFL 1, "
                       "x1.5 from the initial dt, capped at CFL 1000), Newton rtol 1e-3 (<= 3 its, bt line search), "
                       "FGMRES(60) rtol 1e-3 with ILU(0) of the coloured FD matrix; tangent with the exact BC "
                       "linearisation, complex step with the reference's frozen lambda BCs",
           **runs, "target_reduction": target, "tangent_s_to_target": tt, "complex_step_s_to_target": tc,
           "speedup_at_equal_reduction": (tc / tt) if (tt and tc) else None}
    # the same ratio at every new low of the complex step's history (the
    # single figure above depends on how long the complex step takes to
    # squeeze out its last, smallest reduction)
    curve, best = [], 1.0
    for red, t in zip(runs["complex_step"]["reduction"][1:], runs["complex_step"]["seconds"][1:]):
        if red < best:
            best = red
            ttan = t_to(runs["tangent"], red)
            curve.append({"reduction": red, "complex_step_s": t, "tangent_s": ttan,
                          "speedup": (t / ttan) if ttan else None})
    out["speedup_curve"] = curve
    # J*v kernel times on the free-stream state
    v = torch.sin(torch.arange(ev.ndofs, device="cuda", dtype=torch.float64))
    o = ev.empty()
    out["jv_kernel_s"] = {"tangent": time_kernel(torch, lambda: ev.jvp(u0, v, 1.0, out=o), 50),
                          "complex_step": time_kernel(torch, lambda: ev.jvp_cstep(u0, v, 1.0, 1e-7, out=o), 50)}
    out["speedup_per_jv"] = out["jv_kernel_s"]["complex_step"] / out["jv_kernel_s"]["tangent"]
    del ev, u0, v, o
    torch.cuda.empty_cache()
    return out


def bench_newton_configured(torch, fr, steps=2):
    """The configured schedule as the reference's harness sets it
    (tools/boundary_conforming_claude.py:64-127: -ts_type pseudo, increment 1.1
    from the initial dt at CFL 10, here capped at CFL 1000; snes rtol 1e-8,
    stol 1e-8, max_it 40; ksp rtol 1e-5; ILU(0) of the coloured FD matrix,
    flexible GMRES(30) bounded at 400 iterations per solve) from the free
    stream, both products, `steps` steps."""
    from fr_sdg_amd import configs as CF
    cfg = CF.bump(tangent_bc_mode=0)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
    opts = dict(precond=3, ksp_pc_side=2, ksp_restart=30, ksp_rtol=1e-5, ksp_max_it=400, snes_rtol=1e-8,
                snes_stol=1e-8, snes_max_it=40)
    out = {"workload": "bump 256x64 p=5, configured schedule from the free stream: CFL 10, x1.1 from the initial dt, "
                       "cap CFL 1000, Newton rtol 1e-8 / stol 1e-8 / 40 its, FGMRES(30) rtol 1e-5 (<= 400 its), "
                       "ILU(0) of the coloured FD matrix"}
    for name, jt in (("tangent", 0), ("complex_step", 1)):
        out[name] = newton_schedule(torch, fr, ev, u0, jt, 10.0, 1.1, 1000.0, steps, 600.0, opts, "configured " + name)
    del ev, u0
    torch.cuda.empty_cache()
    return out


def bench_rows(torch, fr, reps=50):
    """Cost of the widened rows on their own workloads (stream time per launch
    from HIP events over `reps` launches):
    * positivity-preserving log variables (-cond_sys): the IFunction and the
      tangent product on the bump 256x64 p=5 with and without the change of
      variables; the difference is the map (k_p2c / k_p2c_tan), 16 B/DOF;
    * radial coordinates: the residual kernel on 128x128 p=4 quads (y >= 0)
      with and without the source (1/r per node read in the last phase)."""
    from fr_sdg_amd import configs as CF
    s = torch.cuda.current_stream()

    def t_of(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(reps):
            fn()
        e.record(s)
        e.synchronize()
        return a.elapsed_time(e) / reps / 1e3

    out = {}
    res = {}
    for cs in (0, 1):
        cfg = dict(CF.bump(tangent_bc_mode=0), cond_sys=cs)
        ev = fr.PhysicsEvaluator(cfg)
        u = ev.impose_ic(fr.IC_UNIFORM, CF.bump_freestream())
        u.mul_(1.0 + 1e-3 * torch.sin(torch.arange(ev.ndofs, device="cuda", dtype=torch.float64)))
        ud = torch.zeros_like(u)
        v = torch.cos(torch.arange(ev.ndofs, device="cuda", dtype=torch.float64))
        o = ev.empty()
        res[cs] = (t_of(lambda: ev.fun_implicit(u, ud, out=o)), t_of(lambda: ev.jvp(u, v, 1.0, out=o)), ev.ndofs)
        del ev, u, ud, v, o
        torch.cuda.empty_cache()
    nd = res[0][2]
    dif = res[1][0] - res[0][0]
    out["log_variables"] = {"workload": "bump 256x64 p=5 (2.36 M DOFs)", "ifunction_s": res[0][0],
                            "ifunction_cond_sys_s": res[1][0], "jvp_s": res[0][1], "jvp_cond_sys_s": res[1][1],
                            "map_overhead_s": dif, "map_GBps": (16.0 * nd / dif / 1e9) if dif > 0 else None}
    rad = {}
    for cs in (0, 1):
        cfg = dict(dims=2, nx=128, ny=128, p=4, lx=10.0, ly=10.0, riemann=1, bc_type=[1, 1, 1, 1], coord_sys=cs)
        ev = fr.PhysicsEvaluator(cfg)
        u = ev.impose_ic(fr.IC_VORTEX, [1.0, 1.0, 5.0, 1.0, 10.0])
        o = ev.empty()
        rad[cs] = t_of(lambda: ev.residual_kernel(u, out=o))
        del ev, u, o
        torch.cuda.empty_cache()
    out["radial"] = {"workload": "128x128 p=4 quads on [0,10]^2, HLLC (1.64 M DOFs)", "residual_s": rad[0],
                     "residual_radial_s": rad[1], "overhead": rad[1] / rad[0] - 1.0}
    return out


def batch_setup(fr, scale=1.0):
    """The 64-instance radial shock-tube batch (meshes scaled by `scale`) with
    its initial state and per-instance dt."""
    from fr_sdg_amd import configs as CF
    cfgs = CF.batch64(scale=scale)
    b = fr.Batch(cfgs)
    u = b.empty()
    dts = []
    # the tanh layer of the jump keeps its width in elements on scaled meshes
    prm = CF.EXPLOSION_PRM + [CF.EXPLOSION_SMOOTH_PRM[-1] / scale]
    for k, iv in enumerate(b.instances):
        uk = iv.impose_ic(fr.IC_EXPLOSION, prm, out=b.view(u, k))
        h = cfgs[k]["lx"] / cfgs[k]["nx"]
        dts.append(CF.explicit_cfl(cfgs[k]["p"]) * h / CF.sound_speed_max(uk, cfgs[k]))
    b.set_dt(dts)
    return b, u


def bench_batch(torch, fr, steps=20, scale=1.0):
    """64 radial shock tubes (N in {32..192}, p in {2..6}, seeded), SSP-RK3 at
    CFL 0.4/(2p+1) per instance: 3 persistent batched launches per step (graph
    replay, autotuned launch shape) vs one set of launches per instance."""
    from fr_sdg_amd import autotune
    b, u = batch_setup(fr, scale)
    best, table = autotune.tune_batch(b, u, use_cache=False)
    s = torch.cuda.current_stream()

    def timed(fn):
        w = u.clone()
        fn(w, 2)
        torch.cuda.synchronize()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn(w, steps)
        e.record(s)
        e.synchronize()
        return a.elapsed_time(e) / 1e3 / steps, w

    t_b, wb = timed(lambda w, n: b.ssprk_run(w, n))

    def each(w, n):
        for _ in range(n):
            b.ssprk_step_each(w)
    t_e, we = timed(each)
    same = bool(torch.equal(wb, we))
    ok = bool(torch.isfinite(wb).all())
    mesh = "seeded N" if scale == 1.0 else f"seeded N scaled by {scale}"
    return {"workload": f"64 radial shock tubes (tanh-smoothed jump, width {0.1 / scale:g}), {mesh}/p/gas, Rusanov, "
                        "filter s=16 a=36, SSP-RK3 at CFL 0.4/(2p+1)", "finite": ok,
            "ndofs": b.ndofs, "steps": steps, "batched_s_per_step": t_b, "per_instance_s_per_step": t_e,
            "speedup": t_e / t_b, "dof_rhs_per_s_batched": 3 * b.ndofs / t_b, "tuning": list(best),
            "tuning_table": table, "bitwise_equal": same}


def adjoint_large(torch, fr, args):
    """Unsteady adjoint on the large curved configuration: warped n^2 quads,
    p=7, VdW, HLLC + filter, SSP-RK3 for `adj-steps` steps; J = int rho^2 at
    the final time; gradient w.r.t. the initial state with binomial
    checkpointing (<= adj-snapshots stored states) checked against a central
    finite difference along a smooth direction."""
    from fr_sdg_amd import adjoint as A
    from fr_sdg_amd import configs as CF
    cfg = CF.warped(args.adj_n)
    ev = fr.PhysicsEvaluator(cfg)
    u0 = ev.impose_ic(fr.IC_EXPLOSION, CF.WARPED_PRM)
    h = 2.0 / args.adj_n
    dt = CF.explicit_cfl(7) * h / CF.sound_speed_max(u0, cfg)
    T = args.adj_steps
    m = ev.mass()
    print(f"[adjoint-large] n={args.adj_n} steps={T} dt={dt:.3e} ndofs={ev.ndofs}", file=sys.stderr, flush=True)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    adj = A.UnsteadyAdjoint(ev, dt, T, max_snapshots=args.adj_snapshots)
    uT, t_fwd = timed(lambda: adj.run(u0))
    health = ev.step_health()
    J0, g = A.rho_squared(ev, uT, m)
    print(f"[adjoint-large] forward {t_fwd:.1f} s J={J0:.15e}", file=sys.stderr, flush=True)
    adj.forward_steps = 0
    t_start = time.perf_counter()

    def progress(a):
        if a.adjoint_steps % 200 == 0:
            print(f"[adjoint-large] {a.adjoint_steps}/{T} adjoint steps, {a.forward_steps} forward steps, "
                  f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
    adj.on_adjoint_step = progress
    lam, t_adj = timed(lambda: adj.gradient(u0, g))
    print(f"[adjoint-large] adjoint {t_adj:.1f} s, recomputed steps {adj.forward_steps}", file=sys.stderr, flush=True)
    # a smooth direction: derivative of the initial state with respect to the
    # inner pressure of the radial shock tube (central difference of the IC)
    dp = 1e-3
    prm_p, prm_m = list(CF.WARPED_PRM), list(CF.WARPED_PRM)
    prm_p[1] *= 1.0 + dp
    prm_m[1] *= 1.0 - dp
    v = (ev.impose_ic(fr.IC_EXPLOSION, prm_p) - ev.impose_ic(fr.IC_EXPLOSION, prm_m)) / (2.0 * dp * CF.WARPED_PRM[1])
    def J(u):
        return A.rho_squared(ev, A.UnsteadyAdjoint(ev, dt, T).run(u), m)[0]
    ad = float(torch.dot(lam, v))
    fds = []
    for eps in (float(x) for x in args.adj_fd_eps.split(",")):
        (jp, jm), t_fd = timed(lambda: (J(u0 + eps * v), J(u0 - eps * v)))
        fds.append({"eps": eps, "dJ_v": (jp - jm) / (2 * eps), "rel_diff": abs(ad - (jp - jm) / (2 * eps)) /
                    abs((jp - jm) / (2 * eps)), "seconds": t_fd})
        print(f"[adjoint-large] FD eps {eps:g}: {fds[-1]['dJ_v']:.9e} (adjoint {ad:.9e})", file=sys.stderr, flush=True)
    fd, eps, t_fd = fds[0]["dJ_v"], fds[0]["eps"], fds[0]["seconds"]
    return {"workload": f"warped {args.adj_n}^2 p=7 VdW HLLC+filter, SSP-RK3 x {T}, J = int rho^2",
            "ndofs": ev.ndofs, "dt": dt, "snapshots": args.adj_snapshots,
            "repetition_number": A.repetitions(T, args.adj_snapshots),
            "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
            "forward_health_flags": health, "recomputed_forward_steps": adj.forward_steps,
            "peak_snapshots": adj.peak_snapshots, "J": J0, "dJ_v_adjoint": ad, "dJ_v_central_fd": fd,
            "rel_diff": abs(ad - fd) / abs(fd), "fd_eps": eps, "fd_seconds": t_fd, "fd_by_eps": fds}


def cpu_baseline(args, dt):
    """The CPU oracle (C++ restatement of the reference's loops) on the
    headline workload: SSP-RK3 steps (3 residuals + the post-step filter) of
    the same 1024^2 p=7 state, on all host threads (OpenMP over element rows,
    as the reference runs one MPI rank per core) and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from fr_sdg_amd import configs as CF

    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    o = pyoracle.Oracle(CF.warped(1024))
    u = o.impose_ic(2, CF.WARPED_PRM)

    def run(nthreads, min_s, max_steps):
        pyoracle.set_threads(nthreads)
        steps, t0, x = 0, time.perf_counter(), u
        while True:
            x = o.ssprk_step(x, dt)
            steps += 1
            el = time.perf_counter() - t0
            if el >= min_s or steps >= max_steps:
                return steps, el
    steps_all, el_all = run(threads, 8.0, 20)
    # one thread: one residual evaluation of the same state (a whole step is ~45 s there)
    pyoracle.set_threads(1)
    t0 = time.perf_counter()
    o.residual(u)
    el_one = time.perf_counter() - t0
    pyoracle.set_threads(threads)
    nd = o.ndofs
    return {"value": 3 * nd * steps_all / el_all, "unit": "DOF-RHS evals/s", "cores": threads, "kind": "port",
            "sample": f"{steps_all} SSP-RK3 steps (3 residuals + filter each) of the full headline state "
                      f"({nd} DOFs) by the oracle on {threads} threads, {el_all:.1f} s",
            "value_1thread": nd / el_one,
            "sample_1thread": f"1 residual evaluation of the same state on 1 thread, {el_one:.1f} s",
            "host_cpus": os.cpu_count()}


if __name__ == "__main__":
    main()

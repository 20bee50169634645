"""fr_sdg_amd — MI355X-native FR residual (FP64) behind the reference's callback API.

The compute path is libfrsdg.so (hand-written HIP kernels for gfx950, C-ABI
in include/frsdg.h).  This module is a thin ctypes binding over that ABI with
torch tensors as device-memory plumbing:

* ``PhysicsEvaluator`` mirrors the reference's evaluator / PETSc callback set
  (``residual_kernel``, ``rhs_func``, ``fun_implicit``, ``pre_jv_eval`` +
  ``matmult``, ``limit_solution`` / ``apply_stabilization``), with the same
  sign conventions (R = +div F, rhs = -R, F = U_t + R; petsc_solver_impl.hpp:
  626-670) and the reference's natural DMDA layout;
* ``FRResidual`` is a ``torch.autograd.Function``: forward = residual kernel,
  backward = adjoint (J^T w) kernel.

There is no CPU fallback: without a GPU or without libfrsdg.so every call
raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must load torch's HIP runtime before libfrsdg.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrsdg.so")

# boundary types (boundary_conditions.hpp:8-16) + the curvilinear slip wall
BC_WALL, BC_OUTFLOW, BC_DIRICHLET, BC_PERIODIC, BC_SLIPWALL = 0, 1, 2, 5, 6
# directions (boundary_conditions.hpp:17-23)
LEFT, RIGHT, TOP, BOTTOM = 0, 1, 2, 3
GEOM_CARTESIAN, GEOM_DELAVAL, GEOM_WEDGE, GEOM_BUMP, GEOM_WARPED = 0, 1, 2, 3, 4
RUSANOV, HLLC = 0, 1
IDEAL, VAN_DER_WAALS = 0, 1
IC_UNIFORM, IC_VORTEX, IC_EXPLOSION, IC_SOD, IC_PULSE = 0, 1, 2, 3, 4


class FrsdgError(RuntimeError):
    pass


class FrsdgDesc(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("p", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("gas_kind", ctypes.c_int), ("gamma", ctypes.c_double), ("R_spec", ctypes.c_double),
        ("vdw_a", ctypes.c_double), ("vdw_b", ctypes.c_double),
        ("riemann", ctypes.c_int), ("geometry", ctypes.c_int),
        ("geom_param", ctypes.c_double * 4),
        ("bc_type", ctypes.c_int * 4), ("bc_state", (ctypes.c_double * 4) * 4),
        ("c_vcjh", ctypes.c_double),
        ("apply_filter", ctypes.c_int), ("filt_cutoff", ctypes.c_int),
        ("filt_strength", ctypes.c_int), ("filt_alpha", ctypes.c_double),
        ("jac_type", ctypes.c_int), ("tangent_bc_mode", ctypes.c_int),
        ("apply_weno", ctypes.c_int), ("weno_gamma", ctypes.c_double),
        ("shock_theta", ctypes.c_double), ("shock_threshold", ctypes.c_double),
        ("cond_sys", ctypes.c_int), ("coord_sys", ctypes.c_int),
    ]


class NewtonOpts(ctypes.Structure):
    _fields_ = [("jac_type", ctypes.c_int), ("precond", ctypes.c_int), ("snes_rtol", ctypes.c_double),
                ("snes_atol", ctypes.c_double), ("snes_max_it", ctypes.c_int), ("ksp_restart", ctypes.c_int),
                ("ksp_max_it", ctypes.c_int), ("ksp_rtol", ctypes.c_double), ("ksp_cgs_refine", ctypes.c_int),
                ("snes_stol", ctypes.c_double), ("ksp_pc_side", ctypes.c_int), ("pc_lag", ctypes.c_int),
                ("pc_fp32", ctypes.c_int)]


class NewtonStats(ctypes.Structure):
    _fields_ = [("newton_its", ctypes.c_int), ("ksp_its", ctypes.c_int), ("jv_evals", ctypes.c_int),
                ("res_evals", ctypes.c_int), ("converged", ctypes.c_int), ("fnorm0", ctypes.c_double),
                ("fnorm", ctypes.c_double), ("seconds", ctypes.c_double), ("pc_builds", ctypes.c_int),
                ("t_pc_assemble", ctypes.c_double), ("t_pc_factor", ctypes.c_double),
                ("t_pc_apply", ctypes.c_double), ("t_jv", ctypes.c_double), ("t_orth", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdjointStats(ctypes.Structure):
    _fields_ = [("forward_steps", ctypes.c_int), ("adjoint_steps", ctypes.c_int), ("snapshots_used", ctypes.c_int),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# PETSc defaults: SNES rtol 1e-8, atol 1e-50, stol 1e-8, max_it 50; KSP GMRES restart 30, rtol 1e-5,
# max_it 10000; preconditioner rebuilt every Newton iteration.  ksp_pc_side stays right (0) here, as the
# residual-history checks assume; PETSc's GMRES default is left (1).
NEWTON_DEFAULTS = dict(jac_type=0, precond=0, snes_rtol=1e-8, snes_atol=1e-50, snes_max_it=50, ksp_restart=30,
                       ksp_max_it=10000, ksp_rtol=1e-5, ksp_cgs_refine=0, snes_stol=1e-8, ksp_pc_side=0, pc_lag=1,
                       pc_fp32=0)

EXPORTS = [
    "frsdg_create", "frsdg_destroy", "frsdg_ndofs", "frsdg_get_tables", "frsdg_residual",
    "frsdg_rhs", "frsdg_ifunction", "frsdg_jvp", "frsdg_jvp_cstep", "frsdg_vjp",
    "frsdg_pre_jv_eval", "frsdg_matmult", "frsdg_filter", "frsdg_apply_stabilization",
    "frsdg_ssprk_step", "frsdg_ssprk_run", "frsdg_impose_ic", "frsdg_l2_error", "frsdg_strerror",
    "frsdg_beuler_step", "frsdg_batch_create", "frsdg_batch_destroy", "frsdg_batch_ndofs",
    "frsdg_batch_offsets", "frsdg_batch_instance", "frsdg_batch_set_dt", "frsdg_batch_set_tuning",
    "frsdg_batch_ssprk_step", "frsdg_batch_ssprk_run", "frsdg_batch_ssprk_step_each", "frsdg_batch_set_tile_shape",
    "frsdg_filter_transpose", "frsdg_mass", "frsdg_boundary_weights", "frsdg_adjoint_solve", "frsdg_pseudo_run",
    "frsdg_color_jacobian", "frsdg_ilu0_factor", "frsdg_ilu0_apply", "frsdg_linear_solve",
    "frsdg_ilu0_apply_transpose", "frsdg_l2_error_noalloc", "frsdg_unsteady_adjoint", "frsdg_step_health",
    "frsdg_check_state", "frsdg_set_stop_mask", "frsdg_batch_health",
]

# health flags of stored states (frsdg_step_health / frsdg_check_state)
HEALTH_NONFINITE, HEALTH_DENSITY, HEALTH_ENERGY = 1, 2, 4

_lib = None


def load_library(path=LIB_PATH):
    """Load libfrsdg.so (built in-tree by ``fr_sdg_amd.build``); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FrsdgError(f"libfrsdg.so not built ({path}); run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    vp, dp, d, i = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.frsdg_create.argtypes = [ctypes.POINTER(FrsdgDesc), ctypes.POINTER(ctypes.c_void_p)]
    L.frsdg_destroy.argtypes = [vp]
    L.frsdg_ndofs.argtypes = [vp]
    L.frsdg_ndofs.restype = ctypes.c_longlong
    L.frsdg_get_tables.argtypes = [vp] + [ctypes.POINTER(ctypes.c_double)] * 6
    L.frsdg_residual.argtypes = [vp, d, dp, dp, vp]
    L.frsdg_rhs.argtypes = [vp, d, dp, dp, vp]
    L.frsdg_ifunction.argtypes = [vp, d, dp, dp, dp, vp]
    L.frsdg_jvp.argtypes = [vp, d, dp, dp, d, dp, vp]
    L.frsdg_jvp_cstep.argtypes = [vp, d, dp, dp, d, d, dp, vp]
    L.frsdg_vjp.argtypes = [vp, d, dp, dp, d, dp, vp]
    L.frsdg_pre_jv_eval.argtypes = [vp, d, dp, d]
    L.frsdg_matmult.argtypes = [vp, dp, dp, vp]
    L.frsdg_filter.argtypes = [vp, dp, dp, vp]
    L.frsdg_apply_stabilization.argtypes = [vp, dp, vp]
    L.frsdg_ssprk_step.argtypes = [vp, d, d, dp, i, vp]
    L.frsdg_ssprk_run.argtypes = [vp, d, d, dp, i, i, vp]
    L.frsdg_impose_ic.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double), dp, vp]
    L.frsdg_l2_error.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double), d, dp,
                                 ctypes.POINTER(ctypes.c_double), vp]
    L.frsdg_step_health.argtypes = [vp, ctypes.POINTER(ctypes.c_int), i, vp]
    L.frsdg_check_state.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int), vp]
    L.frsdg_set_stop_mask.argtypes = [vp, i]
    L.frsdg_unsteady_adjoint.argtypes = [vp, d, d, i, i, dp, dp, dp, ctypes.POINTER(AdjointStats), vp]
    L.frsdg_l2_error_noalloc.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double), d, dp,
                                         ctypes.POINTER(ctypes.c_double), vp]
    L.frsdg_beuler_step.argtypes = [vp, d, d, dp, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats), vp]
    L.frsdg_pseudo_run.argtypes = [vp, d, d, d, i, d, d, i, dp, ctypes.POINTER(NewtonOpts),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(NewtonStats), vp]
    L.frsdg_batch_create.argtypes = [ctypes.POINTER(FrsdgDesc), i, ctypes.POINTER(ctypes.c_void_p)]
    L.frsdg_batch_destroy.argtypes = [vp]
    L.frsdg_batch_ndofs.argtypes = [vp]
    L.frsdg_batch_ndofs.restype = ctypes.c_longlong
    L.frsdg_batch_offsets.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.frsdg_batch_instance.argtypes = [vp, i, ctypes.POINTER(ctypes.c_void_p)]
    L.frsdg_batch_set_dt.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.frsdg_batch_set_tuning.argtypes = [vp, i, i, i]
    L.frsdg_batch_set_tile_shape.argtypes = [vp, i, i]
    L.frsdg_batch_ssprk_step.argtypes = [vp, dp, vp]
    L.frsdg_batch_ssprk_run.argtypes = [vp, dp, i, vp]
    L.frsdg_batch_health.argtypes = [vp, ctypes.POINTER(ctypes.c_int), i, vp]
    L.frsdg_batch_ssprk_step_each.argtypes = [vp, ctypes.POINTER(ctypes.c_double), dp, vp]
    L.frsdg_filter_transpose.argtypes = [vp, dp, dp, vp]
    L.frsdg_mass.argtypes = [vp, dp, vp]
    L.frsdg_boundary_weights.argtypes = [vp, i, dp, vp]
    L.frsdg_color_jacobian.argtypes = [vp, d, dp, dp, d, i, dp, vp]
    L.frsdg_linear_solve.argtypes = [vp, d, d, dp, dp, dp, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats),
                                     ctypes.POINTER(ctypes.c_double), i, ctypes.POINTER(ctypes.c_int), vp]
    L.frsdg_ilu0_factor.argtypes = [vp, dp, dp, vp]
    L.frsdg_ilu0_apply.argtypes = [vp, dp, dp, dp, dp, vp]
    L.frsdg_ilu0_apply_transpose.argtypes = [vp, dp, dp, dp, dp, vp]
    L.frsdg_adjoint_solve.argtypes = [vp, d, d, dp, dp, dp, ctypes.POINTER(NewtonOpts),
                                      ctypes.POINTER(NewtonStats), vp]
    L.frsdg_strerror.argtypes = [i]
    L.frsdg_strerror.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = _lib.frsdg_strerror(rc).decode() if _lib is not None else str(rc)
        raise FrsdgError(f"libfrsdg error {rc}: {msg}")


def make_desc(cfg):
    """Descriptor from a plain config dict (keys = frsdg_desc field names)."""
    d = FrsdgDesc()
    d.dims = cfg.get("dims", 2)
    d.nx = cfg["nx"]
    d.ny = cfg.get("ny", 1)
    d.p = cfg["p"]
    d.lx = cfg.get("lx", 1.0)
    d.ly = cfg.get("ly", 1.0)
    d.gas_kind = cfg.get("gas_kind", IDEAL)
    d.gamma = cfg.get("gamma", 1.4)
    d.R_spec = cfg.get("R_spec", 287.05)
    d.vdw_a = cfg.get("vdw_a", 0.0)
    d.vdw_b = cfg.get("vdw_b", 0.0)
    d.riemann = cfg.get("riemann", HLLC)
    d.geometry = cfg.get("geometry", GEOM_CARTESIAN)
    for k, v in enumerate(cfg.get("geom_param", [0.0] * 4)):
        d.geom_param[k] = v
    for k, v in enumerate(cfg.get("bc_type", [BC_OUTFLOW] * 4)):
        d.bc_type[k] = v
    for k, st in enumerate(cfg.get("bc_state", [[0.0] * 4] * 4)):
        for q, v in enumerate(st):
            d.bc_state[k][q] = v
    d.c_vcjh = cfg.get("c_vcjh", 0.0)
    d.apply_filter = cfg.get("apply_filter", 0)
    d.filt_cutoff = cfg.get("filt_cutoff", 2)
    d.filt_strength = cfg.get("filt_strength", 32)
    d.filt_alpha = cfg.get("filt_alpha", 36.0)
    d.jac_type = cfg.get("jac_type", 0)
    d.tangent_bc_mode = cfg.get("tangent_bc_mode", 0)
    d.apply_weno = cfg.get("apply_weno", 0)
    d.weno_gamma = cfg.get("weno_gamma", 0.998)  # parser.hpp WenoConstants defaults
    d.shock_theta = cfg.get("shock_theta", 1.0)
    d.shock_threshold = cfg.get("shock_threshold", 0.8)
    d.cond_sys = cfg.get("cond_sys", 0)
    d.coord_sys = cfg.get("coord_sys", 0)
    return d


def _ptr(t):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise FrsdgError("expected a contiguous float64 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class PhysicsEvaluator:
    """Device-resident FR evaluator for one problem (physics_evaluator.hpp:158-318)."""

    def __init__(self, cfg, device="cuda"):
        if not torch.cuda.is_available():
            raise FrsdgError("fr_sdg_amd needs a ROCm GPU (no CPU fallback)")
        self.lib = load_library()
        self.cfg = dict(cfg)
        self.device = torch.device(device)
        torch.cuda.init()
        self._desc = make_desc(cfg)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(self.lib.frsdg_create(ctypes.byref(self._desc), ctypes.byref(h)))
        self.h = h
        self.n = cfg["p"] + 1
        self.ndofs = int(self.lib.frsdg_ndofs(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.frsdg_destroy(self.h)
            self.h = None

    __del__ = close

    # -- helpers --
    def empty(self):
        return torch.empty(self.ndofs, dtype=torch.float64, device=self.device)

    def tables(self):
        import numpy as np
        n = self.n
        arrs = [np.zeros(n), np.zeros(n), np.zeros(n * n), np.zeros(n), np.zeros(n), np.zeros(n * n)]
        _check(self.lib.frsdg_get_tables(self.h, *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                   for a in arrs]))
        csi, w, D, dL, dR, F = arrs
        return dict(csi=csi, w=w, D=D.reshape(n, n), dL=dL, dR=dR, filt=F.reshape(n, n))

    # -- residual callbacks --
    def residual_kernel(self, u, t=0.0, out=None):
        """R = +div F (physics_impl_2D.hpp:644-902)."""
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_residual(self.h, t, _ptr(u), _ptr(out), _stream()))
        return out

    def rhs_func(self, u, t=0.0, out=None):
        """rhs = -R(u) (petsc_solver_impl.hpp:657-670)."""
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_rhs(self.h, t, _ptr(u), _ptr(out), _stream()))
        return out

    def fun_implicit(self, u, udot, t=0.0, out=None):
        """F = U_t + R(u) (petsc_solver_impl.hpp:626-655)."""
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_ifunction(self.h, t, _ptr(u), _ptr(udot), _ptr(out), _stream()))
        return out

    # -- Jacobian products --
    def jvp(self, u, v, sigma=0.0, t=0.0, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_jvp(self.h, t, _ptr(u), _ptr(v), sigma, _ptr(out), _stream()))
        return out

    def jvp_cstep(self, u, v, sigma=0.0, h=1e-7, t=0.0, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_jvp_cstep(self.h, t, _ptr(u), _ptr(v), sigma, h, _ptr(out), _stream()))
        return out

    def vjp(self, u, w, sigma=0.0, t=0.0, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_vjp(self.h, t, _ptr(u), _ptr(w), sigma, _ptr(out), _stream()))
        return out

    def pre_jv_eval(self, u, sigma, t=0.0):
        """Snapshot of the linearisation point (petsc_solver_impl.hpp:223-248)."""
        self._jv_u = u  # keep the tensor alive: the library holds its pointer
        _check(self.lib.frsdg_pre_jv_eval(self.h, t, _ptr(u), sigma))

    def matmult(self, v, out=None):
        """MatShell MATOP_MULT (petsc_solver_impl.hpp:250-267)."""
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_matmult(self.h, _ptr(v), _ptr(out), _stream()))
        return out

    # -- stabilisation / stepping --
    def limit_solution(self, u, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_filter(self.h, _ptr(u), _ptr(out), _stream()))
        return out

    def apply_stabilization(self, u):
        _check(self.lib.frsdg_apply_stabilization(self.h, _ptr(u), _stream()))
        return u

    def ssprk_step(self, u, dt, t=0.0, stages=3):
        _check(self.lib.frsdg_ssprk_step(self.h, t, dt, _ptr(u), stages, _stream()))
        return u

    def ssprk_run(self, u, dt, nsteps, t=0.0, stages=3):
        """nsteps graph-replayed steps; raises FrsdgError (numerical failure) when a
        stage output goes non-finite or non-positive (the run stops there)."""
        _check(self.lib.frsdg_ssprk_run(self.h, t, dt, _ptr(u), stages, nsteps, _stream()))
        return u

    def step_health(self, clear=True):
        """Health flags accumulated by ssprk_step (HEALTH_* bits; 0 = healthy)."""
        f = ctypes.c_int(0)
        _check(self.lib.frsdg_step_health(self.h, ctypes.byref(f), int(clear), _stream()))
        return f.value

    def set_stop_mask(self, mask):
        """HEALTH_* bits that stop ssprk_run with a numerical failure (default
        HEALTH_NONFINITE)."""
        _check(self.lib.frsdg_set_stop_mask(self.h, int(mask)))

    def check_state(self, u):
        """Health flags of a conservative state (HEALTH_* bits; 0 = healthy)."""
        f = ctypes.c_int(0)
        _check(self.lib.frsdg_check_state(self.h, _ptr(u), ctypes.byref(f), _stream()))
        return f.value

    def beuler_step(self, u, dt, t=0.0, **opts):
        """One backward-Euler step solved by JFNK (TSSetIFunction path); u updated in place."""
        o = dict(NEWTON_DEFAULTS)
        o.update(opts)
        no = NewtonOpts(**o)
        st = NewtonStats()
        _check(self.lib.frsdg_beuler_step(self.h, t, dt, _ptr(u), ctypes.byref(no), ctypes.byref(st), _stream()))
        return st.as_dict()

    def pseudo_run(self, u, dt0, increment, max_steps, from_initial_dt=True, dt_max=0.0, t_max=0.0, t=0.0,
                   **opts):
        """Pseudo-transient continuation (TSPSEUDO, TSPseudoTimeStepDefault);
        u updated in place.  Returns the step sizes, the steady-residual norms
        after each step and the cumulative wall seconds."""
        o = dict(NEWTON_DEFAULTS)
        o.update(opts)
        no = NewtonOpts(**o)
        dts = (ctypes.c_double * max(1, max_steps))()
        fns = (ctypes.c_double * max(1, max_steps))()
        tms = (ctypes.c_double * max(1, max_steps))()
        n = ctypes.c_int(0)
        st = NewtonStats()
        _check(self.lib.frsdg_pseudo_run(self.h, t, dt0, increment, int(from_initial_dt), dt_max, t_max, max_steps,
                                         _ptr(u), ctypes.byref(no), dts, fns, tms, ctypes.byref(n), ctypes.byref(st),
                                         _stream()))
        return {"dt": list(dts[:n.value]), "fnorm": list(fns[:n.value]), "seconds": list(tms[:n.value]),
                "steps": n.value, "fnorm0": st.fnorm0, "stats": st.as_dict()}

    def color_layout(self):
        """(E, S, nd): block rows, slots per row (9 in 2D, 3 in 1D), DOFs per element."""
        d = self._desc
        E = d.nx * (d.ny if d.dims == 2 else 1)
        return E, (9 if d.dims == 2 else 3), self.ndofs // E

    def color_jacobian(self, u, sigma, u_n=None, method=0, t=0.0):
        """The reference's preconditioning matrix dF/du (DMDA box-stencil
        colouring + MatFDColoringApply, petsc_solver_impl.hpp:72-95,126-160)
        of F(x) = IFunction(t, x, sigma (T(x) - T(u_n))).  Returns a tensor
        [E, S, nd(col), nd(row)]: block (e, slot) = dF_e/du_neighbour.
        method 0 FD with PETSc's default "wp" step, 1 "ds" step, 2 exact."""
        E, S, nd = self.color_layout()
        P = torch.empty((E, S, nd, nd), dtype=torch.float64, device=self.device)
        _check(self.lib.frsdg_color_jacobian(self.h, t, _ptr(u), None if u_n is None else _ptr(u_n), sigma,
                                             int(method), _ptr(P), _stream()))
        return P

    def ilu0_factor(self, P):
        """Block ILU(0) of P (color_jacobian layout) in place; returns the
        diagonal blocks' triangular inverses [E, 2, nd, nd]."""
        E, S, nd = self.color_layout()
        Dinv = torch.empty((E, 2, nd, nd), dtype=torch.float64, device=self.device)
        _check(self.lib.frsdg_ilu0_factor(self.h, _ptr(P), _ptr(Dinv), _stream()))
        return Dinv

    def ilu0_apply(self, P, Dinv, x, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_ilu0_apply(self.h, _ptr(P), _ptr(Dinv), _ptr(x), _ptr(out), _stream()))
        return out

    def ilu0_apply_transpose(self, P, Dinv, x, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_ilu0_apply_transpose(self.h, _ptr(P), _ptr(Dinv), _ptr(x), _ptr(out), _stream()))
        return out

    def linear_solve(self, u, b, sigma, t=0.0, out=None, **opts):
        """(sigma I + J(u)) x = b by the device GMRES; returns x, the stats and
        the residual history (||b||, then the estimate after each iteration)."""
        out = self.empty() if out is None else out
        o = dict(NEWTON_DEFAULTS)
        o.update(opts)
        no = NewtonOpts(**o)
        st = NewtonStats()
        cap = int(no.ksp_max_it) + 1
        hist = (ctypes.c_double * cap)()
        n = ctypes.c_int(0)
        _check(self.lib.frsdg_linear_solve(self.h, t, sigma, _ptr(u), _ptr(b), _ptr(out), ctypes.byref(no),
                                           ctypes.byref(st), hist, cap, ctypes.byref(n), _stream()))
        return out, st.as_dict(), list(hist[:n.value])

    def adjoint_solve(self, u, rhs, sigma, t=0.0, out=None, **opts):
        """(sigma I + J(u))^T lambda = rhs by GMRES on the adjoint kernel."""
        out = self.empty() if out is None else out
        o = dict(NEWTON_DEFAULTS)
        o.update(opts)
        no = NewtonOpts(**o)
        st = NewtonStats()
        _check(self.lib.frsdg_adjoint_solve(self.h, t, sigma, _ptr(u), _ptr(rhs), _ptr(out), ctypes.byref(no),
                                            ctypes.byref(st), _stream()))
        return out, st.as_dict()

    def unsteady_adjoint(self, u0, lam_T, dt, nsteps, max_snapshots=32, t=0.0, out=None):
        """lambda_0 = (d u_N / d u_0)^T lambda_T over nsteps SSP-RK3 steps (binomial
        checkpointing in the library); returns lambda_0 and the statistics."""
        out = self.empty() if out is None else out
        st = AdjointStats()
        _check(self.lib.frsdg_unsteady_adjoint(self.h, t, dt, int(nsteps), int(max_snapshots), _ptr(u0), _ptr(lam_T),
                                               _ptr(out), ctypes.byref(st), _stream()))
        return out, st.as_dict()

    def filter_transpose(self, w, out=None):
        out = self.empty() if out is None else out
        _check(self.lib.frsdg_filter_transpose(self.h, _ptr(w), _ptr(out), _stream()))
        return out

    def mass(self):
        m = self.empty()
        _check(self.lib.frsdg_mass(self.h, _ptr(m), _stream()))
        return m

    def boundary_weights(self, direction):
        along = self.cfg["ny"] if direction < 2 else self.cfg["nx"]
        out = torch.empty(along * self.n * 2, dtype=torch.float64, device=self.device)
        _check(self.lib.frsdg_boundary_weights(self.h, direction, _ptr(out), _stream()))
        return out.view(along, self.n, 2)

    def impose_ic(self, kind, prm, out=None):
        out = self.empty() if out is None else out
        prm = (ctypes.c_double * 8)(*(list(prm) + [0.0] * (8 - len(prm))))
        _check(self.lib.frsdg_impose_ic(self.h, kind, prm, _ptr(out), _stream()))
        return out

    def compute_error(self, kind, prm, u, t=0.0):
        prm = (ctypes.c_double * 8)(*(list(prm) + [0.0] * (8 - len(prm))))
        out = (ctypes.c_double * 4)()
        _check(self.lib.frsdg_l2_error(self.h, kind, prm, t, _ptr(u), out, _stream()))
        return list(out)


    def compute_error_noalloc(self, kind, prm, u, t=0.0):
        """compute_error through the allocation-free entry point (static device
        scratch, at most 2^24 nodes; calls that overlap on one GPU race)."""
        prm = (ctypes.c_double * 8)(*(list(prm) + [0.0] * (8 - len(prm))))
        out = (ctypes.c_double * 4)()
        _check(self.lib.frsdg_l2_error_noalloc(self.h, kind, prm, t, _ptr(u), out, _stream()))
        return list(out)


class _InstanceView(PhysicsEvaluator):
    """Non-owning evaluator over one instance of a Batch."""

    def __init__(self, lib, h, cfg, device):
        self.lib, self.h, self.cfg, self.device = lib, h, dict(cfg), device
        self.n = cfg["p"] + 1
        self.ndofs = int(lib.frsdg_ndofs(h))

    def close(self):
        self.h = None

    __del__ = close


class Batch:
    """Many independent problems advanced by persistent batched kernels."""

    def __init__(self, cfgs, device="cuda"):
        if not torch.cuda.is_available():
            raise FrsdgError("fr_sdg_amd needs a ROCm GPU (no CPU fallback)")
        self.lib = load_library()
        self.cfgs = [dict(c) for c in cfgs]
        self.device = torch.device(device)
        torch.cuda.init()
        arr = (FrsdgDesc * len(cfgs))(*[make_desc(c) for c in cfgs])
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(self.lib.frsdg_batch_create(arr, len(cfgs), ctypes.byref(h)))
        self.h = h
        self.count = len(cfgs)
        self.ndofs = int(self.lib.frsdg_batch_ndofs(h))
        off = (ctypes.c_longlong * (self.count + 1))()
        _check(self.lib.frsdg_batch_offsets(h, off))
        self.offsets = list(off)
        self.instances = []
        for k in range(self.count):
            ck = ctypes.c_void_p()
            _check(self.lib.frsdg_batch_instance(h, k, ctypes.byref(ck)))
            self.instances.append(_InstanceView(self.lib, ck, cfgs[k], self.device))

    def close(self):
        if getattr(self, "h", None):
            for iv in self.instances:
                iv.close()
            self.lib.frsdg_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def empty(self):
        return torch.empty(self.ndofs, dtype=torch.float64, device=self.device)

    def view(self, u, k):
        return u[self.offsets[k]:self.offsets[k + 1]]

    def set_dt(self, dts):
        self.dts = list(dts)
        _check(self.lib.frsdg_batch_set_dt(self.h, (ctypes.c_double * self.count)(*self.dts)))

    def set_tuning(self, grid, epb_div, mode=1):
        _check(self.lib.frsdg_batch_set_tuning(self.h, grid, epb_div, mode))

    def set_tile_shape(self, n, shape):
        """Element tile of the order-n instances: 0 default, 1 halved, 2 the
        other tile with the fewest rim faces per element."""
        _check(self.lib.frsdg_batch_set_tile_shape(self.h, n, shape))

    def orders(self):
        return sorted({c["p"] + 1 for c in self.cfgs})

    def ssprk_step(self, u):
        _check(self.lib.frsdg_batch_ssprk_step(self.h, _ptr(u), _stream()))
        return u

    def ssprk_run(self, u, nsteps):
        _check(self.lib.frsdg_batch_ssprk_run(self.h, _ptr(u), nsteps, _stream()))
        return u

    def health(self, clear=False):
        """Health flags of each instance's stage outputs since they were last
        cleared (bit 1 non-finite, bit 2 rho <= 0, bit 4 internal energy <= 0)."""
        f = (ctypes.c_int * self.count)()
        _check(self.lib.frsdg_batch_health(self.h, f, int(clear), _stream()))
        return list(f)

    def ssprk_step_each(self, u):
        _check(self.lib.frsdg_batch_ssprk_step_each(self.h, (ctypes.c_double * self.count)(*self.dts),
                                                    _ptr(u), _stream()))
        return u


class FRResidual(torch.autograd.Function):
    """R(u) with the adjoint kernel as backward: grad_u = J(u)^T grad_R."""

    @staticmethod
    def forward(ctx, u, ev, t=0.0):
        ctx.ev, ctx.t = ev, t
        ctx.save_for_backward(u)
        return ev.residual_kernel(u.detach().contiguous(), t)

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        return ctx.ev.vjp(u.detach().contiguous(), g.contiguous(), 0.0, ctx.t), None, None

    @staticmethod
    def jvp(ctx, du, _ev=None, _t=None):
        (u,) = ctx.saved_tensors
        return ctx.ev.jvp(u.detach().contiguous(), du.contiguous(), 0.0, ctx.t)


def fr_residual(u, ev, t=0.0):
    return FRResidual.apply(u, ev, t)

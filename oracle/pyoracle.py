"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, and only as the checker.  The product (``fr_sdg_amd``) never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class FroDesc(ctypes.Structure):
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
        ("apply_weno", ctypes.c_int), ("weno_gamma", ctypes.c_double),
        ("shock_theta", ctypes.c_double), ("shock_threshold", ctypes.c_double),
        ("cond_sys", ctypes.c_int), ("coord_sys", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.fro_create.restype = vp
        L.fro_create.argtypes = [ctypes.POINTER(FroDesc)]
        L.fro_set_threads.argtypes = [ctypes.c_int]
        L.fro_set_threads.restype = ctypes.c_int
        L.fro_destroy.argtypes = [vp]
        L.fro_ndofs.restype = ctypes.c_long
        L.fro_ndofs.argtypes = [vp]
        L.fro_tables.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.fro_residual.argtypes = [vp, ctypes.c_double, dp, dp]
        L.fro_residual_cplx.argtypes = [vp, ctypes.c_double, dp, dp, dp, dp]
        L.fro_jvp_cstep.argtypes = [vp, ctypes.c_double, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.fro_filter.argtypes = [vp, dp, dp]
        L.fro_ifunction.argtypes = [vp, ctypes.c_double, dp, dp, dp]
        L.fro_prim_to_cons.argtypes = [vp, dp, dp]
        L.fro_cons_to_prim.argtypes = [vp, dp, dp]
        L.fro_ssprk_step.argtypes = [vp, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int]
        L.fro_impose_ic.argtypes = [vp, ctypes.c_int, dp, dp]
        L.fro_l2_error.argtypes = [vp, ctypes.c_int, dp, ctypes.c_double, dp, dp]
        L.fro_flux2d.argtypes = [vp, dp, dp, dp]
        L.fro_riemann2d.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_double, dp]
        L.fro_riemann1d.argtypes = [vp, dp, dp, dp]
        L.fro_metrics.argtypes = [vp, ctypes.c_double, ctypes.c_double, dp]
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the oracle's OpenMP element loops (results do not depend on
    it); returns the previous setting."""
    return lib().fro_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_desc(cfg):
    """Build an oracle descriptor from a plain config dict (same keys as the product)."""
    d = FroDesc()
    d.dims = cfg.get("dims", 2)
    d.nx = cfg["nx"]
    d.ny = cfg.get("ny", 1)
    d.p = cfg["p"]
    d.lx = cfg.get("lx", 1.0)
    d.ly = cfg.get("ly", 1.0)
    d.gas_kind = cfg.get("gas_kind", 0)
    d.gamma = cfg.get("gamma", 1.4)
    d.R_spec = cfg.get("R_spec", 287.05)
    d.vdw_a = cfg.get("vdw_a", 0.0)
    d.vdw_b = cfg.get("vdw_b", 0.0)
    d.riemann = cfg.get("riemann", 1)
    d.geometry = cfg.get("geometry", 0)
    for i, v in enumerate(cfg.get("geom_param", [0.0] * 4)):
        d.geom_param[i] = v
    for i, v in enumerate(cfg.get("bc_type", [1, 1, 1, 1])):
        d.bc_type[i] = v
    for i, st in enumerate(cfg.get("bc_state", [[0.0] * 4] * 4)):
        for q, v in enumerate(st):
            d.bc_state[i][q] = v
    d.c_vcjh = cfg.get("c_vcjh", 0.0)
    d.apply_filter = cfg.get("apply_filter", 0)
    d.filt_cutoff = cfg.get("filt_cutoff", 2)
    d.filt_strength = cfg.get("filt_strength", 32)
    d.filt_alpha = cfg.get("filt_alpha", 36.0)
    d.apply_weno = cfg.get("apply_weno", 0)
    d.weno_gamma = cfg.get("weno_gamma", 0.998)
    d.shock_theta = cfg.get("shock_theta", 1.0)
    d.shock_threshold = cfg.get("shock_threshold", 0.8)
    d.cond_sys = cfg.get("cond_sys", 0)
    d.coord_sys = cfg.get("coord_sys", 0)
    return d


class Oracle:
    """CPU restatement of the reference evaluator for one configuration."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        self._d = make_desc(cfg)
        self.h = lib().fro_create(ctypes.byref(self._d))
        if not self.h:
            raise RuntimeError("oracle: invalid configuration")
        self.n = cfg["p"] + 1
        self.ndofs = lib().fro_ndofs(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().fro_destroy(self.h)
            self.h = None

    def tables(self):
        n = self.n
        out = [np.zeros(n), np.zeros(n), np.zeros(n * n), np.zeros(n), np.zeros(n), np.zeros(n * n)]
        lib().fro_tables(self.h, *[_p(a) for a in out])
        csi, w, D, dL, dR, F = out
        return dict(csi=csi, w=w, D=D.reshape(n, n), dL=dL, dR=dR, filt=F.reshape(n, n))

    def residual(self, u, t=0.0):
        u = _c(u)
        R = np.zeros(self.ndofs)
        lib().fro_residual(self.h, t, _p(u), _p(R))
        return R

    def residual_cplx(self, u, t=0.0):
        u = np.asarray(u, dtype=np.complex128)
        ure, uim = _c(u.real), _c(u.imag)
        Rr, Ri = np.zeros(self.ndofs), np.zeros(self.ndofs)
        lib().fro_residual_cplx(self.h, t, _p(ure), _p(uim), _p(Rr), _p(Ri))
        return Rr + 1j * Ri

    def jvp_cstep(self, u, v, sigma=0.0, h=1e-7, t=0.0):
        u, v = _c(u), _c(v)
        out = np.zeros(self.ndofs)
        lib().fro_jvp_cstep(self.h, t, _p(u), _p(v), sigma, h, _p(out))
        return out

    def ifunction(self, u, udot, t=0.0):
        u, udot = _c(u), _c(udot)
        F = np.zeros(self.ndofs)
        lib().fro_ifunction(self.h, t, _p(u), _p(udot), _p(F))
        return F

    def prim_to_cons(self, w):
        w = _c(w)
        c = np.zeros(self.ndofs)
        lib().fro_prim_to_cons(self.h, _p(w), _p(c))
        return c

    def cons_to_prim(self, c):
        c = _c(c)
        w = np.zeros(self.ndofs)
        lib().fro_cons_to_prim(self.h, _p(c), _p(w))
        return w

    def filter(self, u):
        u = _c(u)
        out = np.zeros(self.ndofs)
        lib().fro_filter(self.h, _p(u), _p(out))
        return out

    def ssprk_step(self, u, dt, t=0.0, stages=3):
        u = _c(u).copy()
        rc = lib().fro_ssprk_step(self.h, t, dt, _p(u), stages)
        if rc:
            raise ValueError("bad stage count")
        return u

    def impose_ic(self, kind, prm):
        prm = _c(list(prm) + [0.0] * (8 - len(prm)))
        u = np.zeros(self.ndofs)
        lib().fro_impose_ic(self.h, kind, _p(prm), _p(u))
        return u

    def l2_error(self, kind, prm, u, t=0.0):
        prm = _c(list(prm) + [0.0] * (8 - len(prm)))
        u = _c(u)
        out = np.zeros(4)
        lib().fro_l2_error(self.h, kind, _p(prm), t, _p(u), _p(out))
        return out

    def flux2d(self, U):
        U = _c(U)
        F, G = np.zeros(4), np.zeros(4)
        lib().fro_flux2d(self.h, _p(U), _p(F), _p(G))
        return F, G

    def riemann2d(self, UL, UR, nx, ny):
        UL, UR = _c(UL), _c(UR)
        out = np.zeros(4)
        lib().fro_riemann2d(self.h, _p(UL), _p(UR), nx, ny, _p(out))
        return out

    def riemann1d(self, UL, UR):
        UL, UR = _c(UL), _c(UR)
        out = np.zeros(3)
        lib().fro_riemann1d(self.h, _p(UL), _p(UR), _p(out))
        return out

    def metrics(self, cx, cy):
        out = np.zeros(9)
        lib().fro_metrics(self.h, cx, cy, _p(out))
        return out

"""Second, independent restatement of the 2D FR residual in vectorised numpy
(test infrastructure).  It is written from the mathematics of the method, not
from the oracle's loops, and is meant for small meshes:

* tables: GLL nodes as the roots of (1 - x^2) P_p'(x) (numpy Legendre series),
  D = V' V^-1 with the Legendre Vandermonde matrices, and the VCJH correction
  derivatives g_L', g_R' of Vincent, Castonguay & Jameson (2011):
      g_R = 1/2 [P_p + (eta P_{p-1} + P_{p+1}) / (1 + eta)],
      g_L = (-1)^p/2 [P_p - (eta P_{p-1} + P_{p+1}) / (1 + eta)],
      eta = c (2p+1) (a_p p!)^2 / 2,  a_p = (2p)! / (2^p (p!)^2);
* fluxes and the ideal / van der Waals closures;
* Rusanov and HLLC in Toro's standard form  F*_K = F_K + s_K (U*_K - U_K)
  (the reference writes the star flux in the equivalent p* form,
  riemann_solver.hpp:249-349), with the reference's smoothed wave-speed bounds
  (smooth_abs with eps = 1e-7, macro_and_libc.hpp:80-113);
* chain-rule volume term R = xi_x dF/dxi + eta_x dF/deta + xi_y dG/dxi +
  eta_y dG/deta at every node (physics_impl_2D.hpp:267-357), every face flux
  evaluated once and lifted into both neighbours with dR (left/bottom
  element) and dL (right/top element) (physics_impl_2D.hpp:689-899);
* boundary states: outflow, wall, Dirichlet, curved slip wall, periodic wrap.

Supported maps: Cartesian, and the sine-warped square (geometry 4) with the
face normal grad(xi)/|grad(xi)| and scale |grad(xi)|.
"""
import math

import numpy as np
from numpy.polynomial import legendre as L

EPS_SMOOTH = 1e-7


def tables(p, c=0.0):
    n = p + 1
    cp = np.zeros(p + 1)
    cp[p] = 1.0
    inner = np.sort(np.real(L.legroots(L.legder(cp)))) if p > 1 else np.array([])
    x = np.concatenate([[-1.0], inner, [1.0]])
    V = np.stack([L.legval(x, np.eye(n)[j]) for j in range(n)], axis=1)
    Vd = np.stack([L.legval(x, L.legder(np.eye(n)[j])) for j in range(n)], axis=1)
    D = np.linalg.solve(V.T, Vd.T).T  # D = Vd V^-1
    ap = math.factorial(2 * p) / (2 ** p * math.factorial(p) ** 2)
    eta = c * (2 * p + 1) * (ap * math.factorial(p)) ** 2 / 2.0

    def leg(k):
        e = np.zeros(p + 2)
        e[k] = 1.0
        return e
    corr = (eta * leg(p - 1) + leg(p + 1)) / (1.0 + eta)
    gR = 0.5 * (leg(p) + corr)
    gL = (-1) ** p * 0.5 * (leg(p) - corr)
    dR = L.legval(x, L.legder(gR))
    dL = L.legval(x, L.legder(gL))
    return x, D, dL, dR


def prims(U, gas):
    rho = np.maximum(U[..., 0], np.finfo(float).tiny)
    u = U[..., 1] / rho
    v = U[..., 2] / rho
    E = np.maximum(U[..., 3], np.finfo(float).tiny)
    e = E / rho - 0.5 * (u * u + v * v)
    g = gas["gamma"]
    if gas["kind"] == 0:
        p = (g - 1.0) * rho * e
        c = np.sqrt(g * p / rho)
    else:
        a, b = gas["a"], gas["b"]
        cov = 1.0 - b * rho
        p = (g - 1.0) * rho * (e + a * rho) / cov - a * rho * rho
        c = np.sqrt(g * (g - 1.0) * (e + a * rho) / cov ** 2 + (g - 1.0) * a * rho * (2.0 - b * rho) / cov ** 2
                    - 2.0 * a * rho)
    return rho, u, v, p, c


def fluxes(U, gas):
    rho, u, v, p, _ = prims(U, gas)
    F = np.stack([U[..., 1], U[..., 1] * u + p, U[..., 1] * v, u * (U[..., 3] + p)], axis=-1)
    G = np.stack([U[..., 2], U[..., 1] * v, U[..., 2] * v + p, v * (U[..., 3] + p)], axis=-1)
    return F, G


def sabs(x):
    return np.sqrt(x * x + EPS_SMOOTH ** 2) - EPS_SMOOTH


def smax(a, b):
    return 0.5 * (a + b + sabs(a - b))


def smin(a, b):
    return 0.5 * (a + b - sabs(a - b))


def riemann(UL, UR, nx, ny, gas, kind):
    """Numerical flux through a face with unit normal (nx, ny), physical frame."""
    _, uL, vL, pL, cL = prims(UL, gas)
    _, uR, vR, pR, cR = prims(UR, gas)
    rL, rR = UL[..., 0], UR[..., 0]
    unL, utL = uL * nx + vL * ny, -uL * ny + vL * nx
    unR, utR = uR * nx + vR * ny, -uR * ny + vR * nx
    # rotated conserved states and fluxes
    QL = np.stack([rL, rL * unL, rL * utL, UL[..., 3]], axis=-1)
    QR = np.stack([rR, rR * unR, rR * utR, UR[..., 3]], axis=-1)
    FL = np.stack([rL * unL, rL * unL * unL + pL, rL * unL * utL, unL * (UL[..., 3] + pL)], axis=-1)
    FR = np.stack([rR * unR, rR * unR * unR + pR, rR * unR * utR, unR * (UR[..., 3] + pR)], axis=-1)
    if kind == 0:
        s = smax(sabs(unL) + cL, sabs(unR) + cR)[..., None]
        f = 0.5 * (FL + FR) - 0.5 * s * (QR - QL)
    else:
        sL = smin(unL - cL, unR - cR)
        sR = smax(unL + cL, unR + cR)
        sM = (pR - pL + rL * unL * (sL - unL) - rR * unR * (sR - unR)) / (rL * (sL - unL) - rR * (sR - unR))

        def star(r, un, ut, E, p, sK):
            fac = r * (sK - un) / (sK - sM)
            return np.stack([fac, fac * sM, fac * ut, fac * (E / r + (sM - un) * (sM + p / (r * (sK - un))))],
                            axis=-1)
        FsL = FL + sL[..., None] * (star(rL, unL, utL, UL[..., 3], pL, sL) - QL)
        FsR = FR + sR[..., None] * (star(rR, unR, utR, UR[..., 3], pR, sR) - QR)
        f = np.where((sL >= 0)[..., None], FL,
                     np.where((sM >= 0)[..., None], FsL, np.where((sR >= 0)[..., None], FsR, FR)))
    return np.stack([f[..., 0], f[..., 1] * nx - f[..., 2] * ny, f[..., 1] * ny + f[..., 2] * nx, f[..., 3]],
                    axis=-1)


def mapping(cfg, cx, cy):
    """Physical-map derivatives (x_xi, x_eta, y_xi, y_eta) at composite coordinates."""
    Lx, Ly = cfg["lx"], cfg.get("ly", 1.0)
    geom = cfg.get("geometry", 0)
    one, zero = np.ones_like(cx), np.zeros_like(cx)
    if geom == 0:
        return Lx * one, zero, zero, Ly * one
    assert geom == 4, "numpy restatement covers the Cartesian and warped maps"
    A, x0, y0 = cfg["geom_param"][0], cfg["geom_param"][1], cfg["geom_param"][2]
    X, Y = x0 + Lx * cx, y0 + Ly * cy
    sx, cxx = np.sin(np.pi * X), np.cos(np.pi * X)
    sy, cyy = np.sin(np.pi * Y), np.cos(np.pi * Y)
    dX = A * np.pi * cxx * sy * Lx
    dY = A * np.pi * sx * cyy * Ly
    return Lx + dX, dY, dX, Ly + dY


def ref_metrics(cfg, cx, cy):
    """(xi_x, xi_y, eta_x, eta_y) of the reference-element coordinates (r, s)."""
    xa, xb, ya, yb = mapping(cfg, cx, cy)
    hr, hs = 0.5 / cfg["nx"], 0.5 / cfg["ny"]
    xr, xs, yr, ys = xa * hr, xb * hs, ya * hr, yb * hs
    det = xr * ys - xs * yr
    return ys / det, -xs / det, -yr / det, xr / det


def boundary_state(cfg, d, Uin, nx, ny):
    t = cfg["bc_type"][d]
    if t == 1:
        return Uin.copy()
    if t == 0:
        ex = Uin.copy()
        ex[..., 1 if d < 2 else 2] *= -1.0
        return ex
    if t == 2:
        return np.broadcast_to(np.asarray(cfg["bc_state"][d], dtype=float), Uin.shape).copy()
    if t == 6:
        r = Uin[..., 0]
        u, v = Uin[..., 1] / r, Uin[..., 2] / r
        vn = u * nx + v * ny
        return np.stack([r, r * (u - 2 * vn * nx), r * (v - 2 * vn * ny), Uin[..., 3]], axis=-1)
    raise ValueError(f"boundary type {t}")


def residual(cfg, u):
    p = cfg["p"]
    n = p + 1
    nx, ny = cfg["nx"], cfg["ny"]
    gas = dict(kind=cfg.get("gas_kind", 0), gamma=cfg.get("gamma", 1.4), a=cfg.get("vdw_a", 0.0),
               b=cfg.get("vdw_b", 0.0))
    rk = cfg.get("riemann", 0)
    x, D, dL, dR = tables(p, cfg.get("c_vcjh", 0.0))
    U = np.asarray(u, dtype=float).reshape(ny, nx, n, n, 4)  # [j, i, l, k, q]
    F, G = fluxes(U, gas)
    # volume term
    ii, kk = np.arange(nx), np.arange(n)
    jj, ll = np.arange(ny), np.arange(n)
    cxv = ((ii[:, None] + 0.5 * (1 + x[None, :])) / nx)            # [i, k]
    cyv = ((jj[:, None] + 0.5 * (1 + x[None, :])) / ny)            # [j, l]
    CX = np.broadcast_to(cxv[None, :, None, :], (ny, nx, n, n))
    CY = np.broadcast_to(cyv[:, None, :, None], (ny, nx, n, n))
    rx, ry, sx, sy = ref_metrics(cfg, CX, CY)
    dFr = np.einsum("km,jilmq->jilkq", D, F)
    dGr = np.einsum("km,jilmq->jilkq", D, G)
    dFs = np.einsum("lm,jimkq->jilkq", D, F)
    dGs = np.einsum("lm,jimkq->jilkq", D, G)
    R = (rx[..., None] * dFr + sx[..., None] * dFs) + (ry[..., None] * dGr + sy[..., None] * dGs)
    per_x = cfg["bc_type"][0] == 5 and cfg["bc_type"][1] == 5
    per_y = cfg["bc_type"][2] == 5 and cfg["bc_type"][3] == 5
    # x faces I = 0..nx between elements I-1 and I of each row
    I = np.arange(nx + 1)
    cxf = np.broadcast_to((I / nx)[None, :, None], (ny, nx + 1, n))
    cyf = np.broadcast_to(cyv[:, None, :], (ny, nx + 1, n))
    mrx, mry, _, _ = ref_metrics(cfg, cxf, cyf)
    sc = np.sqrt(mrx * mrx + mry * mry)
    fnx, fny = mrx / sc, mry / sc
    UL = np.empty((ny, nx + 1, n, 4))
    UR = np.empty((ny, nx + 1, n, 4))
    UL[:, 1:] = U[:, :, :, n - 1, :]
    UR[:, :-1] = U[:, :, :, 0, :]
    if per_x:
        UL[:, 0] = U[:, nx - 1, :, n - 1, :]
        UR[:, nx] = U[:, 0, :, 0, :]
    else:
        UL[:, 0] = boundary_state(cfg, 0, UR[:, 0], fnx[:, 0], fny[:, 0])
        UR[:, nx] = boundary_state(cfg, 1, UL[:, nx], fnx[:, nx], fny[:, nx])
    Fs = riemann(UL, UR, fnx, fny, gas, rk)
    FLx, FLy = fluxes(UL, gas)
    FRx, FRy = fluxes(UR, gas)
    dA = sc[..., None] * (Fs - (FLx * fnx[..., None] + FLy * fny[..., None]))  # left element's jump
    dB = sc[..., None] * (Fs - (FRx * fnx[..., None] + FRy * fny[..., None]))  # right element's jump
    R += dA[:, 1:, :, None, :] * dR[None, None, None, :, None]
    R += dB[:, :-1, :, None, :] * dL[None, None, None, :, None]
    # y faces J = 0..ny between elements J-1 and J of each column
    J = np.arange(ny + 1)
    cxf = np.broadcast_to(cxv[None, :, :], (ny + 1, nx, n))
    cyf = np.broadcast_to((J / ny)[:, None, None], (ny + 1, nx, n))
    _, _, msx, msy = ref_metrics(cfg, cxf, cyf)
    sc = np.sqrt(msx * msx + msy * msy)
    fnx, fny = msx / sc, msy / sc
    UB = np.empty((ny + 1, nx, n, 4))
    UT = np.empty((ny + 1, nx, n, 4))
    UB[1:] = U[:, :, n - 1, :, :]
    UT[:-1] = U[:, :, 0, :, :]
    if per_y:
        UB[0] = U[ny - 1, :, n - 1, :, :]
        UT[ny] = U[0, :, 0, :, :]
    else:
        UB[0] = boundary_state(cfg, 3, UT[0], fnx[0], fny[0])
        UT[ny] = boundary_state(cfg, 2, UB[ny], fnx[ny], fny[ny])
    Fs = riemann(UB, UT, fnx, fny, gas, rk)
    FBx, FBy = fluxes(UB, gas)
    FTx, FTy = fluxes(UT, gas)
    dA = sc[..., None] * (Fs - (FBx * fnx[..., None] + FBy * fny[..., None]))
    dB = sc[..., None] * (Fs - (FTx * fnx[..., None] + FTy * fny[..., None]))
    R += dA[1:, :, None, :, :] * dR[None, None, :, None, None]
    R += dB[:-1, :, None, :, :] * dL[None, None, :, None, None]
    return R.reshape(-1)

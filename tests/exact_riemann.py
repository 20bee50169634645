"""Exact solution of the 1D Euler Riemann problem (ideal gas), test helper.

Restates the textbook method (Toro, "Riemann Solvers and Numerical Methods for
Fluid Dynamics", ch. 4) that tools/sod_shock_moc.py:6-125 of the reference
implements: pressure functions f_K for shock / rarefaction branches, star
pressure by bracketed root finding, star velocity, star densities, and
sampling along x/t.  Used as the method-of-characteristics reference for the
Sod full-run checks.
"""
import numpy as np
from scipy.optimize import brentq


def star_state(rL, uL, pL, rR, uR, pR, g=1.4):
    aL, aR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)

    def fK(p, rK, pK, aK):
        if p > pK:
            A = 2.0 / ((g + 1.0) * rK)
            B = (g - 1.0) / (g + 1.0) * pK
            return (p - pK) * np.sqrt(A / (p + B))
        return 2.0 * aK / (g - 1.0) * ((p / pK) ** ((g - 1.0) / (2.0 * g)) - 1.0)

    ps = brentq(lambda p: fK(p, rL, pL, aL) + fK(p, rR, pR, aR) + (uR - uL), 1e-12, 10 * max(pL, pR),
                xtol=1e-15, rtol=1e-15)
    us = 0.5 * (uL + uR + fK(ps, rR, pR, aR) - fK(ps, rL, pL, aL))
    gm = (g - 1) / (g + 1)
    rsL = rL * (ps / pL + gm) / (gm * ps / pL + 1.0) if ps > pL else rL * (ps / pL) ** (1.0 / g)
    rsR = rR * (ps / pR + gm) / (gm * ps / pR + 1.0) if ps > pR else rR * (ps / pR) ** (1.0 / g)
    return ps, us, rsL, rsR


def sample(rL, uL, pL, rR, uR, pR, t, x, x0=0.5, g=1.4):
    """(rho, u, p) at positions x and time t > 0."""
    aL, aR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    ps, us, rsL, rsR = star_state(rL, uL, pL, rR, uR, pR, g)
    asL, asR = np.sqrt(g * ps / rsL), np.sqrt(g * ps / rsR)
    x = np.asarray(x, dtype=float)
    rho, u, p = np.empty_like(x), np.empty_like(x), np.empty_like(x)
    for i, xi in enumerate(x.ravel()):
        S = (xi - x0) / t
        if S < us:
            if ps > pL:
                SL = uL - aL * np.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                st = (rL, uL, pL) if S < SL else (rsL, us, ps)
            else:
                if S < uL - aL:
                    st = (rL, uL, pL)
                elif S > us - asL:
                    st = (rsL, us, ps)
                else:
                    uf = 2.0 / (g + 1.0) * (aL + (g - 1.0) / 2.0 * uL + S)
                    af = uf - S
                    st = (rL * (af / aL) ** (2.0 / (g - 1.0)), uf, pL * (af / aL) ** (2.0 * g / (g - 1.0)))
        else:
            if ps > pR:
                SR = uR + aR * np.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                st = (rR, uR, pR) if S > SR else (rsR, us, ps)
            else:
                if S > uR + aR:
                    st = (rR, uR, pR)
                elif S < us + asR:
                    st = (rsR, us, ps)
                else:
                    uf = 2.0 / (g + 1.0) * (-aR + (g - 1.0) / 2.0 * uR + S)
                    af = S - uf
                    st = (rR * (af / aR) ** (2.0 / (g - 1.0)), uf, pR * (af / aR) ** (2.0 * g / (g - 1.0)))
        rho.flat[i], u.flat[i], p.flat[i] = st
    return rho, u, p

// frsdg_physics.hpp — pointwise physics shared by every kernel, templated on
// the scalar type (double / Dual / Cplx, frsdg_scalar.hpp).
//
// Formulas follow the reference so that results agree to rounding:
//   smooth_abs/max/min          macro_and_libc.hpp:80-113 (eps = 1e-7)
//   wrap_positive               macro_and_libc.hpp:115-138
//   ideal gas / van der Waals   gas_models.hpp:56-158
//   2D fluxes                   riemann_solver.hpp:48-77
//   rotated Rusanov / HLLC      riemann_solver.hpp:189-349
//   1D flux / Rusanov / HLLC    riemann_solver.hpp:39-46, 87-184
// Arithmetic is regrouped where that saves divides on the GPU (one
// reciprocal of rho per state, one of (sL-sM) per HLLC star state); the
// differences are at rounding level.
#pragma once
#include "frsdg_scalar.hpp"

namespace frsdg
{
    struct GasP
    {
        int    kind; // 0 ideal, 1 van der Waals
        double gamma, a, b;
    };

    template <class S>
    FR_HD S smooth_abs(S x)
    {
        const double eps = 1e-7;
        return sqrt_(x * x + eps * eps) - eps;
    }
    template <class S>
    FR_HD S smooth_max(S a, S b)
    {
        return 0.5 * (a + b + smooth_abs(a - b));
    }
    template <class S>
    FR_HD S smooth_min(S a, S b)
    {
        return 0.5 * (a + b - smooth_abs(a - b));
    }

    // pressure from density and specific internal energy.  Van der Waals:
    // the reciprocal of the co-volume factor 1 - b rho is formed once per
    // state; the sound speed below forms the same one (one reciprocal
    // after common-subexpression elimination where both are evaluated)
    template <class S>
    FR_HD S gas_pressure(const GasP& g, S rho, S e)
    {
        if (g.kind == 0) return (g.gamma - 1.0) * rho * e;
        const S rcov = rcp_(1.0 - g.b * rho);
        const S cvT  = e + g.a * rho;
        return ((g.gamma - 1.0) * rho * cvT) * rcov - g.a * rho * rho;
    }
    // sound speed (ideal: sqrt(gamma p / rho) with p passed in)
    template <class S>
    FR_HD S gas_sound(const GasP& g, S rho, S e, S p, S inv_rho)
    {
        if (g.kind == 0) return sqrt_(g.gamma * p * inv_rho);
        const S rcov  = rcp_(1.0 - g.b * rho);
        const S inv_d = rcov * rcov;
        const S th    = e + g.a * rho;
        const S t1    = (g.gamma * (g.gamma - 1.0) * th) * inv_d;
        const S t2    = ((g.gamma - 1.0) * g.a * rho * (2.0 - g.b * rho)) * inv_d;
        const S t3    = 2.0 * g.a * rho;
        return sqrt_(t1 + t2 - t3);
    }

    // primitive state of a 2D node: floored rho, 1/rho, u, v, e (make_state), p
    template <class S>
    struct Prim2
    {
        S rho, inv_rho, u, v, e, p;
    };
    template <class S>
    FR_HD Prim2<S> prim2(const GasP& g, const S U[4])
    {
        Prim2<S> q;
        q.rho        = Sc<S>::wrap_positive(U[0]);
        q.inv_rho    = rcp_(q.rho);
        q.u          = U[1] * q.inv_rho;
        q.v          = U[2] * q.inv_rho;
        const S totE = Sc<S>::wrap_positive(U[3]);
        q.e          = totE * q.inv_rho - 0.5 * (q.u * q.u + q.v * q.v);
        q.p          = gas_pressure(g, q.rho, q.e);
        return q;
    }
    // euler_fluxes_2d_w_p: F, G (and p) at a node
    template <class S>
    FR_HD void flux2p(const GasP& g, const S U[4], S F[4], S G[4], S& p)
    {
        const Prim2<S> q  = prim2(g, U);
        p                 = q.p;
        const S        Ep = U[3] + q.p;
        const S        uv = U[1] * q.v;
        F[0]              = U[1];
        F[1]              = U[1] * q.u + q.p;
        F[2]              = uv;
        F[3]              = q.u * Ep;
        G[0]              = U[2];
        G[1]              = uv;
        G[2]              = U[2] * q.v + q.p;
        G[3]              = q.v * Ep;
    }
    template <class S>
    FR_HD void flux2(const GasP& g, const S U[4], S F[4], S G[4])
    {
        S p;
        flux2p(g, U, F, G, p);
    }

    // radial (axisymmetric) source of a 2D node, physics_impl_2D.hpp:306-319:
    // G / r for every equation, then -p / r on the radial momentum (ir = 1/r,
    // 0 on the axis)
    template <class S>
    FR_HD void radial_source2(const GasP& g, const S U[4], double ir, S src[4])
    {
        S F[4], G[4], p;
        flux2p(g, U, F, G, p);
#pragma unroll
        for (int q = 0; q < 4; ++q) src[q] = G[q] * ir;
        src[2] = src[2] - p * ir;
    }
    // radial source of a 1D node, physics_impl_1D.hpp:70-90 (no density floor)
    template <class S>
    FR_HD void radial_source1(const GasP& g, const S U[3], double ir, S src[3])
    {
        const S vel  = U[1] / U[0];
        const S pres = gas_pressure(g, U[0], U[2] / U[0] - 0.5 * vel * vel);
        src[0]       = ir * U[1];
        src[1]       = ir * U[1] * vel;
        src[2]       = ir * vel * (U[2] + pres);
    }

    // rotated 2D numerical flux; kind 0 Rusanov, 1 HLLC
    template <class S>
    FR_HD void riemann2(const GasP& g, int kind, const S UL[4], const S UR[4], double nx, double ny,
                        S out[4])
    {
        const Prim2<S> L   = prim2(g, UL);
        const Prim2<S> R   = prim2(g, UR);
        const S        unL = L.u * nx + L.v * ny;
        const S        utL = -L.u * ny + L.v * nx;
        const S        unR = R.u * nx + R.v * ny;
        const S        utR = -R.u * ny + R.v * nx;
        const S        cL  = gas_sound(g, L.rho, L.e, L.p, L.inv_rho);
        const S        cR  = gas_sound(g, R.rho, R.e, R.p, R.inv_rho);
        // normal-frame physical fluxes
        const S mL = UL[0] * unL, mR = UR[0] * unR;
        S       FL[4] = {mL, mL * unL + L.p, mL * utL, (UL[3] + L.p) * unL};
        S       FR[4] = {mR, mR * unR + R.p, mR * utR, (UR[3] + R.p) * unR};
        S       fs[4];
        if (kind == 0)
        {
            const S smax = smooth_max(smooth_abs(unL) + cL, smooth_abs(unR) + cR);
            fs[0]        = 0.5 * (FL[0] + FR[0]) - 0.5 * smax * (UR[0] - UL[0]);
            fs[1]        = 0.5 * (FL[1] + FR[1]) - 0.5 * smax * (UR[0] * unR - UL[0] * unL);
            fs[2]        = 0.5 * (FL[2] + FR[2]) - 0.5 * smax * (UR[0] * utR - UL[0] * utL);
            fs[3]        = 0.5 * (FL[3] + FR[3]) - 0.5 * smax * (UR[3] - UL[3]);
        }
        else
        {
            const S sL  = smooth_min(unL - cL, unR - cR);
            const S sR  = smooth_max(unL + cL, unR + cR);
            const S rLd = UL[0] * (sL - unL);
            const S rRd = UR[0] * (sR - unR);
            const S sM  = quot_(R.p - L.p + unL * rLd - unR * rRd, rLd - rRd);
            // Branch order of the reference: sL >= 0 -> F_L; sM >= 0 -> left
            // star state; sR >= 0 -> right star state; else F_R.  The star
            // state of the selected side is formed once with selects (no
            // divergent double evaluation across the lanes of a wavefront).
            const bool supL = Sc<S>::branch(sL) >= 0.0;
            const bool left = Sc<S>::branch(sM) >= 0.0;
            const bool supR = !left && !(Sc<S>::branch(sR) >= 0.0);
            const S    sK   = Sc<S>::sel(left, sL, sR);
            const S    unK  = Sc<S>::sel(left, unL, unR);
            const S    utK  = Sc<S>::sel(left, utL, utR);
            const S    rK   = Sc<S>::sel(left, UL[0], UR[0]);
            const S    mK   = Sc<S>::sel(left, mL, mR);
            const S    EK   = Sc<S>::sel(left, UL[3], UR[3]);
            const S    pK   = Sc<S>::sel(left, L.p, R.p);
            S          FK[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) FK[q] = Sc<S>::sel(left, FL[q], FR[q]);
            const S ps  = pK + rK * (sK - unK) * (sM - unK);
            const S inv = rcp_(sK - sM);
            S       st[4];
            st[0] = (sM * (sK * rK - FK[0])) * inv;
            st[1] = (sM * (sK * mK - FK[1]) + sK * ps) * inv;
            st[2] = utK * st[0];
            st[3] = (sM * (sK * EK - FK[3]) + sK * ps * sM) * inv;
#pragma unroll
            for (int q = 0; q < 4; ++q) fs[q] = Sc<S>::sel(supL, FL[q], Sc<S>::sel(supR, FR[q], st[q]));
        }
        out[0] = fs[0];
        out[1] = fs[1] * nx - fs[2] * ny;
        out[2] = fs[1] * ny + fs[2] * nx;
        out[3] = fs[3];
    }

    // ---------------- 1D (riemann_solver.hpp:19-46, 87-184) ----------------
    template <class S>
    FR_HD void flux1(const GasP& g, const S U[3], S F[3])
    {
        const S rho  = Sc<S>::wrap_positive(U[0]);
        const S ir   = 1.0 / rho;
        const S u    = U[1] * ir;
        const S totE = Sc<S>::wrap_positive(U[2]);
        const S e    = totE * ir - 0.5 * u * u;
        const S p    = gas_pressure(g, rho, e);
        F[0]         = U[1];
        F[1]         = U[1] * u + p;
        F[2]         = (U[2] + p) * u;
    }
    template <class S>
    FR_HD void riemann1(const GasP& g, int kind, const S UL[3], const S UR[3], S out[3])
    {
        S       FL[3], FR[3];
        flux1(g, UL, FL);
        flux1(g, UR, FR);
        // make_state (floored) for the sound speeds; velocity with the unfloored
        // density as in the reference 1D solvers (quirk: riemann_solver.hpp:102,131)
        const S rhoL = Sc<S>::wrap_positive(UL[0]), rhoR = Sc<S>::wrap_positive(UR[0]);
        const S irL = 1.0 / rhoL, irR = 1.0 / rhoR;
        const S uLf = UL[1] * irL, uRf = UR[1] * irR;
        const S eL  = Sc<S>::wrap_positive(UL[2]) * irL - 0.5 * uLf * uLf;
        const S eR  = Sc<S>::wrap_positive(UR[2]) * irR - 0.5 * uRf * uRf;
        const S pL_ = gas_pressure(g, rhoL, eL), pR_ = gas_pressure(g, rhoR, eR);
        const S aL  = gas_sound(g, rhoL, eL, pL_, irL);
        const S aR  = gas_sound(g, rhoR, eR, pR_, irR);
        const S uL  = UL[1] / UL[0];
        const S uR  = UR[1] / UR[0];
        if (kind == 0)
        {
            const S sL   = smooth_abs(uL) + aL;
            const S sR   = smooth_abs(uR) + aR;
            const S smax = smooth_max(sL, sR);
#pragma unroll
            for (int q = 0; q < 3; ++q) out[q] = 0.5 * (FL[q] + FR[q]) - 0.5 * smax * (UR[q] - UL[q]);
            return;
        }
        const S pL  = FL[1] - uL * UL[1];
        const S pR  = FR[1] - uR * UR[1];
        const S sL  = smooth_min(uL - aL, uR - aR);
        const S sR  = smooth_max(uL + aL, uR + aR);
        const S rLd = UL[0] * (sL - uL);
        const S rRd = UR[0] * (sR - uR);
        const S sM  = (pR - pL + uL * rLd - uR * rRd) / (rLd - rRd);
        if (Sc<S>::branch(sL) >= 0.0)
        {
#pragma unroll
            for (int q = 0; q < 3; ++q) out[q] = FL[q];
        }
        else if (Sc<S>::branch(sM) >= 0.0)
        {
            const S ps  = pL + UL[0] * (sL - uL) * (sM - uL);
            const S inv = 1.0 / (sL - sM);
            out[0]      = (sM * (sL * UL[0] - FL[0])) * inv;
            out[1]      = (sM * (sL * UL[1] - FL[1]) + sL * ps) * inv;
            out[2]      = (sM * (sL * UL[2] - FL[2]) + sL * ps * sM) * inv;
        }
        else if (Sc<S>::branch(sR) >= 0.0)
        {
            const S ps  = pR + UR[0] * (sR - uR) * (sM - uR);
            const S inv = 1.0 / (sR - sM);
            out[0]      = (sM * (sR * UR[0] - FR[0])) * inv;
            out[1]      = (sM * (sR * UR[1] - FR[1]) + sR * ps) * inv;
            out[2]      = (sM * (sR * UR[2] - FR[2]) + sR * ps * sM) * inv;
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 3; ++q) out[q] = FR[q];
        }
    }

    // ---------------- boundary states (boundary_conditions.hpp:214-324) ----------------
    struct BCP
    {
        int    type[4];     // Left, Right, Top, Bottom
        double state[4][4]; // Dirichlet states
        int    freeze;      // 1: lambda-bound BCs carry no derivative (complex step semantics)
    };
    enum : int
    {
        BC_WALL = 0,
        BC_OUTFLOW = 1,
        BC_DIRICHLET = 2,
        BC_PERIODIC = 5,
        BC_SLIPWALL = 6
    };
    // static-index selection (keeps the kernel-argument copy of BCP out of
    // scratch memory: no run-time indexing into it)
    FR_HD int bc_type_of(const BCP& bc, int dir)
    {
        return dir == 0 ? bc.type[0] : dir == 1 ? bc.type[1] : dir == 2 ? bc.type[2] : bc.type[3];
    }
    FR_HD double bc_state_of(const BCP& bc, int dir, int q)
    {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (k == dir && m == q) r = bc.state[k][m];
        return r;
    }
    // exterior state for boundary `dir` (0 L, 1 R, 2 T, 3 B) given the interior
    // trace and the face normal.
    template <class S>
    FR_HD void bc_exterior(const BCP& bc, int dir, const S in[4], double nx, double ny, S ex[4])
    {
        // each case sets the four scalars; ex is written once at the end (per-case
        // stores to ex get merged by the optimiser into one store at a
        // case-dependent index, which puts ex in scratch memory)
        const int t  = bc_type_of(bc, dir);
        S         e0 = in[0], e1 = in[1], e2 = in[2], e3 = in[3];
        if (t == BC_WALL)
        {
            // :300-308, axis = dir / 2
            const double sx = dir < 2 ? -1.0 : 1.0;
            e1              = in[1] * sx;
            e2              = in[2] * (0.0 - sx);
        }
        else if (t == BC_DIRICHLET)
        {
            e0 = Sc<S>::c(bc_state_of(bc, dir, 0));
            e1 = Sc<S>::c(bc_state_of(bc, dir, 1));
            e2 = Sc<S>::c(bc_state_of(bc, dir, 2));
            e3 = Sc<S>::c(bc_state_of(bc, dir, 3));
        }
        else if (t == BC_SLIPWALL)
        {
            // tests/bump.cpp:250-272: reflect the velocity about the face normal
            const S r  = bc.freeze ? Sc<S>::frozen(in[0]) : in[0];
            const S mu = bc.freeze ? Sc<S>::frozen(in[1]) : in[1];
            const S mv = bc.freeze ? Sc<S>::frozen(in[2]) : in[2];
            const S en = bc.freeze ? Sc<S>::frozen(in[3]) : in[3];
            const S uu = mu / r;
            const S vv = mv / r;
            const S vn = uu * nx + vv * ny;
            e0         = r;
            e1         = r * (uu - 2.0 * vn * nx);
            e2         = r * (vv - 2.0 * vn * ny);
            e3         = en;
        }
        ex[0] = e0;
        ex[1] = e1;
        ex[2] = e2;
        ex[3] = e3;
    }
    template <class S>
    FR_HD void bc_exterior1(const BCP& bc, int side, const S in[3], S ex[3])
    {
        const int t = bc_type_of(bc, side);
#pragma unroll
        for (int q = 0; q < 3; ++q) ex[q] = in[q];
        if (t == BC_WALL)
            ex[1] = -in[1];
        else if (t == BC_DIRICHLET)
#pragma unroll
            for (int q = 0; q < 3; ++q) ex[q] = Sc<S>::c(bc_state_of(bc, side, q));
    }
} // namespace frsdg

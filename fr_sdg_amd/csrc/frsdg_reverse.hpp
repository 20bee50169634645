// frsdg_reverse.hpp — hand-written reverse-mode (vector-Jacobian) passes of
// the pointwise physics of frsdg_physics.hpp, in double precision:
//   flux2_vjp    Ubar += (dF/dU)^T Fb + (dG/dU)^T Gb
//   riemann2_vjp ULbar += (dF*/dUL)^T fb,  URbar += (dF*/dUR)^T fb
//   bc_vjp       Ubar += (d ex / d U)^T exb   for the exterior state of bc_exterior
// Each is the exact transpose of the forward code path it mirrors (same
// branch decisions, same floors: wrap_positive passes a derivative of 1 above
// DBL_MIN and 0 at the floor), evaluated as one forward sweep that keeps the
// intermediates in registers followed by one backward sweep — about three
// forward evaluations instead of one forward-dual pass per state component.
#pragma once
#include "frsdg_physics.hpp"

namespace frsdg
{
    // forward intermediates of prim2 (+ sound speed when WITH_C) of one state
    struct PrimT
    {
        double U0, U1, U2, U3;    // raw conserved state
        double rho, ir, u, v, E, e, p, c;
        bool   rho_live, E_live;  // above the floors
    };

    template <bool WITH_C>
    FR_HD void prim_fwd(const GasP& g, const double U[4], PrimT& t)
    {
        t.U0       = U[0];
        t.U1       = U[1];
        t.U2       = U[2];
        t.U3       = U[3];
        t.rho_live = DBL_MIN < U[0];
        t.rho      = t.rho_live ? U[0] : DBL_MIN;
        t.ir       = rcp_(t.rho);
        t.u        = U[1] * t.ir;
        t.v        = U[2] * t.ir;
        t.E_live   = DBL_MIN < U[3];
        t.E        = t.E_live ? U[3] : DBL_MIN;
        t.e        = t.E * t.ir - 0.5 * (t.u * t.u + t.v * t.v);
        t.p        = gas_pressure(g, t.rho, t.e);
        t.c        = WITH_C ? gas_sound(g, t.rho, t.e, t.p, t.ir) : 0.0;
    }

    // cotangents of (rho, ir, u, v, e, p[, c]) -> Ubar (accumulated); without
    // WITH_C the sound speed is not part of the sweep (flux2 does not use it,
    // and it may be undefined at states that never reach a face)
    template <bool WITH_C>
    FR_HD void prim_bwd(const GasP& g, const PrimT& t, double rb, double irb, double ub, double vb,
                                             double eb, double pb, double cb, double Ub[4])
    {
        const double gm = g.gamma - 1.0;
        // c; skipped at a zero cotangent, so that a state with c^2 < 0 (whose
        // NaN sound speed only ever reaches a wave-speed comparison that the
        // forward HLLC resolves to F_R) passes no NaN into Ubar
        if (!WITH_C || cb == 0.0)
        {
        }
        else if (g.kind == 0)
        {
            // c = sqrt(gamma p ir)
            const double k = cb * g.gamma * (0.5 * rcp_(t.c));
            pb += k * t.ir;
            irb += k * t.p;
        }
        else
        {
            const double cov  = 1.0 - g.b * t.rho;
            const double icov = rcp_(cov);
            const double invd = icov * icov;
            const double dinv = 2.0 * g.b * invd * icov; // d invd / d rho
            const double th   = t.e + g.a * t.rho;
            const double c2b  = cb * (0.5 * rcp_(t.c));
            eb += c2b * g.gamma * gm * invd;
            rb += c2b * (g.gamma * gm * (g.a * invd + th * dinv) + gm * g.a * (2.0 - 2.0 * g.b * t.rho) * invd +
                         gm * g.a * t.rho * (2.0 - g.b * t.rho) * dinv - 2.0 * g.a);
        }
        // p = P(rho, e)
        if (g.kind == 0)
        {
            rb += pb * gm * t.e;
            eb += pb * gm * t.rho;
        }
        else
        {
            const double icov = rcp_(1.0 - g.b * t.rho);
            const double th   = t.e + g.a * t.rho;
            eb += pb * gm * t.rho * icov;
            rb += pb * (gm * (th + g.a * t.rho) * icov + gm * t.rho * th * g.b * (icov * icov) - 2.0 * g.a * t.rho);
        }
        // e = E ir - (u^2 + v^2)/2
        const double Eb = eb * t.ir;
        irb += eb * t.E;
        ub -= eb * t.u;
        vb -= eb * t.v;
        // u = U1 ir, v = U2 ir
        Ub[1] += ub * t.ir;
        Ub[2] += vb * t.ir;
        irb += ub * t.U1 + vb * t.U2;
        // ir = 1 / rho
        rb -= irb * t.ir * t.ir;
        if (t.rho_live) Ub[0] += rb;
        if (t.E_live) Ub[3] += Eb;
    }

    // Ubar += (dF/dU)^T Fb + (dG/dU)^T Gb   (flux2)
    FR_HD void flux2_vjp(const GasP& g, const double U[4], const double Fb[4], const double Gb[4],
                                              double Ub[4])
    {
        PrimT t;
        prim_fwd<false>(g, U, t);
        const double Ep = U[3] + t.p;
        // F = [U1, U1 u + p, U1 v, u Ep],  G = [U2, U1 v, U2 v + p, v Ep]
        double       ub = Fb[1] * U[1] + Fb[3] * Ep;
        double       vb = Gb[2] * U[2] + Gb[3] * Ep;
        double       pb = Fb[1] + Gb[2];
        const double uvb = Fb[2] + Gb[1];
        const double Epb = Fb[3] * t.u + Gb[3] * t.v;
        Ub[1] += Fb[0] + Fb[1] * t.u + uvb * t.v;
        vb += uvb * U[1];
        Ub[2] += Gb[0] + Gb[2] * t.v;
        Ub[3] += Epb;
        pb += Epb;
        prim_bwd<false>(g, t, 0.0, 0.0, ub, vb, 0.0, pb, 0.0, Ub);
    }

    FR_HD double sabs_d(double x) // d smooth_abs / dx
    {
        const double eps = 1e-7;
        return quot_(x, sqrt_(x * x + eps * eps));
    }

    // ULb += (dF*/dUL)^T fb, URb += (dF*/dUR)^T fb   (riemann2, double)
    FR_HD void riemann2_vjp(const GasP& g, int kind, const double UL[4], const double UR[4],
                                                 double nx, double ny, const double fb_phys[4], double ULb[4],
                                                 double URb[4])
    {
        PrimT L, R;
        prim_fwd<true>(g, UL, L);
        prim_fwd<true>(g, UR, R);
        const double unL = L.u * nx + L.v * ny, utL = -L.u * ny + L.v * nx;
        const double unR = R.u * nx + R.v * ny, utR = -R.u * ny + R.v * nx;
        const double mL = UL[0] * unL, mR = UR[0] * unR;
        // rotate the output cotangent into the face frame
        double fb[4];
        fb[0] = fb_phys[0];
        fb[1] = fb_phys[1] * nx + fb_phys[2] * ny;
        fb[2] = -fb_phys[1] * ny + fb_phys[2] * nx;
        fb[3] = fb_phys[3];
        // cotangents of the face-frame quantities
        double FLb[4] = {0, 0, 0, 0}, FRb[4] = {0, 0, 0, 0};
        double unLb = 0, utLb = 0, unRb = 0, utRb = 0, cLb = 0, cRb = 0, pLb = 0, pRb = 0;
        double U0Lb = 0, U0Rb = 0, U3Lb = 0, U3Rb = 0, mLb = 0, mRb = 0;
        if (kind == 0)
        {
            // fs = (FL + FR)/2 - smax (QR - QL)/2,  Q = [U0, U0 un, U0 ut, U3]
            const double x1 = smooth_abs(unL) + L.c, x2 = smooth_abs(unR) + R.c;
            const double smx = smooth_max(x1, x2);
            const double QL[4] = {UL[0], UL[0] * unL, UL[0] * utL, UL[3]};
            const double QR[4] = {UR[0], UR[0] * unR, UR[0] * utR, UR[3]};
            double       smb = 0.0, QLb[4], QRb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                FLb[q] += 0.5 * fb[q];
                FRb[q] += 0.5 * fb[q];
                smb -= 0.5 * fb[q] * (QR[q] - QL[q]);
                QRb[q] = -0.5 * smx * fb[q];
                QLb[q] = 0.5 * smx * fb[q];
            }
            const double sd  = sabs_d(x1 - x2);
            const double x1b = smb * 0.5 * (1.0 + sd), x2b = smb * 0.5 * (1.0 - sd);
            unLb += x1b * sabs_d(unL);
            cLb += x1b;
            unRb += x2b * sabs_d(unR);
            cRb += x2b;
            U0Lb += QLb[0] + QLb[1] * unL + QLb[2] * utL;
            unLb += QLb[1] * UL[0];
            utLb += QLb[2] * UL[0];
            U3Lb += QLb[3];
            U0Rb += QRb[0] + QRb[1] * unR + QRb[2] * utR;
            unRb += QRb[1] * UR[0];
            utRb += QRb[2] * UR[0];
            U3Rb += QRb[3];
        }
        else
        {
            const double a1 = unL - L.c, a2 = unR - R.c, b1 = unL + L.c, b2 = unR + R.c;
            const double sL  = smooth_min(a1, a2);
            const double sR  = smooth_max(b1, b2);
            const double rLd = UL[0] * (sL - unL);
            const double rRd = UR[0] * (sR - unR);
            const double num = R.p - L.p + unL * rLd - unR * rRd, den = rLd - rRd;
            const double sM  = quot_(num, den);
            const bool   supL = sL >= 0.0;
            const bool   left = sM >= 0.0;
            const bool   supR = !left && !(sR >= 0.0);
            // The wave-speed cotangents are non-zero only for the star state;
            // the supersonic branches (also taken when a sound speed is NaN:
            // every comparison fails and F_R is selected) leave them out.
            double       sLb = 0, sRb = 0, sMb = 0;
            if (supL)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) FLb[q] += fb[q];
            }
            else if (supR)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) FRb[q] += fb[q];
            }
            else
            {
                // star state of side K
                const double sK = left ? sL : sR, unK = left ? unL : unR, utK = left ? utL : utR;
                const double rK = left ? UL[0] : UR[0], mK = left ? mL : mR, EK = left ? UL[3] : UR[3];
                const double pK = left ? L.p : R.p;
                const double FK0 = mK;
                const double FK1 = mK * unK + pK;
                const double FK3 = (EK + pK) * unK;
                const double ps  = pK + rK * (sK - unK) * (sM - unK);
                const double inv = rcp_(sK - sM);
                const double A0 = sK * rK - FK0, A1 = sK * mK - FK1, A3 = sK * EK - FK3;
                const double st0 = (sM * A0) * inv;
                const double B1  = sM * A1 + sK * ps, B3 = sM * A3 + sK * ps * sM;
                // backward
                double       st0b = fb[0] + fb[2] * utK;
                double       utKb = fb[2] * st0;
                const double st1b = fb[1], st3b = fb[3];
                double       invb = st0b * sM * A0 + st1b * B1 + st3b * B3;
                double       sMbK = st0b * A0 * inv + st1b * inv * A1 + st3b * inv * (A3 + sK * ps);
                const double A0b = st0b * sM * inv, A1b = st1b * inv * sM, A3b = st3b * inv * sM;
                double       sKb = st1b * inv * ps + st3b * inv * ps * sM;
                const double psb = st1b * inv * sK + st3b * inv * sK * sM;
                sKb += -invb * inv * inv;
                sMbK += invb * inv * inv;
                // A terms
                sKb += A0b * rK + A1b * mK + A3b * EK;
                double rKb = A0b * sK;
                double mKb = A1b * sK;
                double EKb = A3b * sK;
                const double FK0b = -A0b, FK1b = -A1b, FK3b = -A3b;
                // ps
                double pKb = psb;
                rKb += psb * (sK - unK) * (sM - unK);
                sKb += psb * rK * (sM - unK);
                sMbK += psb * rK * (sK - unK);
                double unKb = -psb * rK * ((sM - unK) + (sK - unK));
                // FK0 = mK, FK1 = mK unK + pK, FK3 = (EK + pK) unK
                mKb += FK0b + FK1b * unK;
                unKb += FK1b * mK + FK3b * (EK + pK);
                pKb += FK1b + FK3b * unK;
                EKb += FK3b * unK;
                sMb += sMbK;
                if (left)
                {
                    sLb += sKb; unLb += unKb; utLb += utKb; U0Lb += rKb; mLb += mKb; U3Lb += EKb; pLb += pKb;
                }
                else
                {
                    sRb += sKb; unRb += unKb; utRb += utKb; U0Rb += rKb; mRb += mKb; U3Rb += EKb; pRb += pKb;
                }
                // sM = num / den
                const double iden = rcp_(den);
                const double numb = sMb * iden, denb = -numb * sM;
                pRb += numb;
                pLb -= numb;
                unLb += numb * rLd;
                unRb -= numb * rRd;
                const double rLdb = numb * unL + denb, rRdb = -numb * unR - denb;
                // rLd = UL0 (sL - unL), rRd = UR0 (sR - unR)
                U0Lb += rLdb * (sL - unL);
                sLb += rLdb * UL[0];
                unLb -= rLdb * UL[0];
                U0Rb += rRdb * (sR - unR);
                sRb += rRdb * UR[0];
                unRb -= rRdb * UR[0];
                // sL = smooth_min(a1, a2), sR = smooth_max(b1, b2)
                const double da = sabs_d(a1 - a2), db = sabs_d(b1 - b2);
                const double a1b = sLb * 0.5 * (1.0 - da), a2b = sLb * 0.5 * (1.0 + da);
                const double b1b = sRb * 0.5 * (1.0 + db), b2b = sRb * 0.5 * (1.0 - db);
                unLb += a1b + b1b;
                cLb += b1b - a1b;
                unRb += a2b + b2b;
                cRb += b2b - a2b;
            }
        }
        // face-frame fluxes F = [m, m un + p, m ut, (U3 + p) un],  m = U0 un
        mLb += FLb[0] + FLb[1] * unL + FLb[2] * utL;
        unLb += FLb[1] * mL + FLb[3] * (UL[3] + L.p);
        pLb += FLb[1] + FLb[3] * unL;
        utLb += FLb[2] * mL;
        U3Lb += FLb[3] * unL;
        mRb += FRb[0] + FRb[1] * unR + FRb[2] * utR;
        unRb += FRb[1] * mR + FRb[3] * (UR[3] + R.p);
        pRb += FRb[1] + FRb[3] * unR;
        utRb += FRb[2] * mR;
        U3Rb += FRb[3] * unR;
        U0Lb += mLb * unL;
        unLb += mLb * UL[0];
        U0Rb += mRb * unR;
        unRb += mRb * UR[0];
        // rotation un = u nx + v ny, ut = -u ny + v nx
        const double uLb = unLb * nx - utLb * ny, vLb = unLb * ny + utLb * nx;
        const double uRb = unRb * nx - utRb * ny, vRb = unRb * ny + utRb * nx;
        ULb[0] += U0Lb;
        ULb[3] += U3Lb;
        URb[0] += U0Rb;
        URb[3] += U3Rb;
        prim_bwd<true>(g, L, 0.0, 0.0, uLb, vLb, 0.0, pLb, cLb, ULb);
        prim_bwd<true>(g, R, 0.0, 0.0, uRb, vRb, 0.0, pRb, cRb, URb);
    }

    // Ub += (d ex / d U)^T exb for ex = bc_exterior(bc, dir, U, nx, ny)
    FR_HD void bc_vjp(const BCP& bc, int dir, const double U[4], double nx, double ny,
                                           const double exb[4], double Ub[4])
    {
        const int t = bc_type_of(bc, dir);
        if (t == BC_WALL)
        {
            const double sx = dir < 2 ? -1.0 : 1.0;
            Ub[0] += exb[0];
            Ub[1] += exb[1] * sx;
            Ub[2] += exb[2] * (0.0 - sx);
            Ub[3] += exb[3];
        }
        else if (t == BC_DIRICHLET)
        {
        }
        else if (t == BC_SLIPWALL)
        {
            if (bc.freeze) return; // lambda-bound: no derivative (complex-step semantics)
            // ex = [r, r (u - 2 vn nx), r (v - 2 vn ny), E], u = U1/r, v = U2/r, vn = u nx + v ny
            const double r = U[0], ir = rcp_(r), u = U[1] * ir, v = U[2] * ir, vn = u * nx + v * ny;
            const double e1 = u - 2.0 * vn * nx, e2 = v - 2.0 * vn * ny;
            double       rb = exb[0] + exb[1] * e1 + exb[2] * e2;
            const double e1b = exb[1] * r, e2b = exb[2] * r;
            double       ub = e1b, vb = e2b;
            const double vnb = -2.0 * (e1b * nx + e2b * ny);
            ub += vnb * nx;
            vb += vnb * ny;
            Ub[1] += ub * ir;
            Ub[2] += vb * ir;
            rb -= (ub * U[1] + vb * U[2]) * (ir * ir);
            Ub[0] += rb;
            Ub[3] += exb[3];
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) Ub[q] += exb[q];
        }
    }
} // namespace frsdg

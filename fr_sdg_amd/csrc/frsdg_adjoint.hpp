// frsdg_adjoint.hpp — adjoint (vector-Jacobian) kernels: out = sigma*w + (dR/du)^T w.
//
// R of element e is   V_e(u_e) + sum over its 4 face copies of  lift_f (Delta_f)
// with Delta_f = s_f (F*(U_L, U_R; n_f) - F(U_own).n_f).  Transposing:
//   phase 1 (k_adj_lift): per element and face copy, the lifted cotangent
//     Wbar_{e,f}(q, idx) = sum_k W_e(q, line(idx, k)) g'(k)      (g' = dL or dR);
//   phase 2 (k_adj2d): per element,
//     volume: Fbar(q,m,l') = sum_k D[k][m] (xi_x W)(q,k,l') + sum_l D[l][l'] (eta_x W)(q,m,l),
//             Gbar likewise with xi_y, eta_y;  ubar += J_F^T Fbar + J_G^T Gbar per node;
//     faces : ubar_own += s_f J_{F*,own}^T (Wbar_{e,f} + Wbar_{e',f'}) - s_f J_{Fn,own}^T Wbar_{e,f},
//             where (e',f') is the neighbour's copy of the same face (its Wbar is read
//             from the phase-1 buffer) and boundary faces chain through the BC map.
// The small Jacobian-transpose products are hand-written reverse sweeps
// (frsdg_reverse.hpp) of the flux / Riemann / BC code the residual uses, with
// the same branch decisions, so the adjoint linearises exactly the discrete
// residual.
// Every output is written once by one thread: deterministic, no atomics.
#pragma once
#include "frsdg_kernels.hpp"
#include "frsdg_reverse.hpp"

namespace frsdg
{
    // Output of the adjoint kernels: out = a w + b (dR/du)^T w [+ c z1] [+ d z2]
    // (the plain VJP is a = sigma, b = 1; the unsteady adjoint fuses the
    // SSP-RK transpose combinations into it)
    struct AdjOut
    {
        double a, b, c, d;
        const double* z1;
        const double* z2;
        __device__ __forceinline__ double at(long long i, double w, double jtw) const
        {
            double o = a * w + b * jtw;
            if (z1) o += c * z1[i];
            if (z2) o += d * z2[i];
            return o;
        }
    };

    constexpr int ADJ_BS = 128;
    template <int N>
    struct AdjTile
    {
        static constexpr int NN  = N * N;
        static constexpr int EPB = (ADJ_BS / NN) > 0 ? (ADJ_BS / NN) : 1;
    };

    // phase 1: Wbar[e][f][idx][q]
    template <int N>
    __global__ __launch_bounds__(256) void k_adj_lift(long long E, const double* __restrict__ tab,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ wb)
    {
        constexpr int NN = N * N;
        const long long t = blockIdx.x * 256LL + threadIdx.x;
        if (t >= E * 4 * N) return;
        const long long e   = t / (4 * N);
        const int       ft  = static_cast<int>(t - e * 4 * N);
        const int       f   = ft / N, idx = ft - (ft / N) * N;
        const double*   g   = tab + NN + ((f == 0 || f == 2) ? 0 : N); // dL or dR
        const double*   we  = w + e * NN * 4;
        double          acc[4] = {0, 0, 0, 0};
        for (int k = 0; k < N; ++k)
        {
            const int node = (f < 2) ? (idx * N + k) : (k * N + idx); // x faces: row l=idx; y faces: column k=idx
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += we[node * 4 + q] * g[k];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) wb[t * 4 + q] = acc[q];
    }

    // J^T c for c attached to F(U).nx + G(U).ny (Fn) and to the Riemann flux w.r.t. the own state
    struct FaceAdjIn
    {
        double uo[4], un[4]; // own / neighbour states (neighbour unused at boundaries)
        double nx, ny;
        int    own_right;    // 1 if the own state is U_R (left/bottom faces), 0 if U_L
        int    bnd, dir;     // boundary face and its BC direction
    };

    // ubar += J_{F*,own}^T cs + J_{Fn,own}^T cn   (cs, cn already scaled), one
    // reverse sweep through the boundary map, the Riemann flux and the flux
    __device__ __forceinline__ void face_adjoint(const GasP& g, int kind, const BCP& bc, const FaceAdjIn& in,
                                                 const double cs[4], const double cn[4], double ub[4])
    {
        double Un[4];
        if (in.bnd)
            bc_exterior(bc, in.dir, in.uo, in.nx, in.ny, Un);
        else
#pragma unroll
            for (int q = 0; q < 4; ++q) Un[q] = in.un[q];
        double ob[4] = {0, 0, 0, 0}, nb[4] = {0, 0, 0, 0};
        if (in.own_right)
            riemann2_vjp(g, kind, Un, in.uo, in.nx, in.ny, cs, nb, ob);
        else
            riemann2_vjp(g, kind, in.uo, Un, in.nx, in.ny, cs, ob, nb);
        if (in.bnd) bc_vjp(bc, in.dir, in.uo, in.nx, in.ny, nb, ob);
        double Fb[4], Gb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            Fb[q] = cn[q] * in.nx;
            Gb[q] = cn[q] * in.ny;
        }
        flux2_vjp(g, in.uo, Fb, Gb, ob);
#pragma unroll
        for (int q = 0; q < 4; ++q) ub[q] += ob[q];
    }
    // only the Riemann part (neighbour copy of a face with its own metrics)
    __device__ __forceinline__ void riemann_adjoint(const GasP& g, int kind, const FaceAdjIn& in,
                                                    const double cs[4], double ub[4])
    {
        double nb[4] = {0, 0, 0, 0};
        if (in.own_right)
            riemann2_vjp(g, kind, in.un, in.uo, in.nx, in.ny, cs, nb, ub);
        else
            riemann2_vjp(g, kind, in.uo, in.un, in.nx, in.ny, cs, ub, nb);
    }

    // phase 2 (2D)
    template <int N>
    __global__ __launch_bounds__(ADJ_BS) void k_adj2d(const Res2DP P, const double* __restrict__ u,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ wb, const AdjOut ao,
                                                       double* __restrict__ out)
    {
        constexpr int NN  = N * N;
        constexpr int EPB = AdjTile<N>::EPB;
        __shared__ double sA[EPB][NN][4], sB[EPB][NN][4], sC[EPB][NN][4], sDd[EPB][NN][4];
        __shared__ double sFc[EPB][4][N][4];
        __shared__ double sD[NN];
        const int tid = threadIdx.x;
        for (int q = tid; q < NN; q += ADJ_BS) sD[q] = P.tab[q];
        const long long blk  = xcd_tile(blockIdx.x, gridDim.x);
        const int       el   = tid / NN;
        const int       node = tid - el * NN;
        const int       k    = node % N;
        const int       l    = node / N;
        const long long e    = blk * EPB + el;
        const bool      act  = (el < EPB) && (e < P.E);
        double W[4], U[4];
        if (act)
        {
            ld4d(w + (e * NN + node) * 4, W);
            ld4d(u + (e * NN + node) * 4, U);
            double mt[4];
            if (P.curved)
                ld4d(P.vmet + (e * NN + node) * 4, mt);
            else
            {
                mt[0] = P.xi_x;
                mt[1] = 0.0;
                mt[2] = 0.0;
                mt[3] = P.eta_y;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                sA[el][node][q]  = mt[0] * W[q]; // xi_x  -> d/dxi  of F
                sB[el][node][q]  = mt[2] * W[q]; // eta_x -> d/deta of F
                sC[el][node][q]  = mt[1] * W[q]; // xi_y  -> d/dxi  of G
                sDd[el][node][q] = mt[3] * W[q]; // eta_y -> d/deta of G
            }
        }
        // ---- face tasks ----
        for (int t = tid; t < EPB * 4 * N; t += ADJ_BS)
        {
            const int       fe = t / (4 * N);
            const int       ft = t - fe * 4 * N;
            const long long eg = blk * EPB + fe;
            if (eg >= P.E) continue;
            const int face = ft / N, idx = ft - (ft / N) * N;
            const int eg32 = static_cast<int>(eg);
            const int j = eg32 / P.nx, i = eg32 - j * P.nx;
            int       ok, ol, nk, nl, ni = i, nj = j, dir, nface;
            bool      bnd;
            if (face == 0)
            {
                ok = 0; ol = idx; nk = N - 1; nl = idx; ni = i - 1; dir = 0; nface = 1;
                bnd = (i == 0) && !P.per_x;
                if (ni < 0) ni += P.nx;
            }
            else if (face == 1)
            {
                ok = N - 1; ol = idx; nk = 0; nl = idx; ni = i + 1; dir = 1; nface = 0;
                bnd = (i == P.nx - 1) && !P.per_x;
                if (ni >= P.nx) ni -= P.nx;
            }
            else if (face == 2)
            {
                ok = idx; ol = 0; nk = idx; nl = N - 1; nj = j - 1; dir = 3; nface = 3;
                bnd = (j == 0) && !P.per_y;
                if (nj < 0) nj += P.ny;
            }
            else
            {
                ok = idx; ol = N - 1; nk = idx; nl = 0; nj = j + 1; dir = 2; nface = 2;
                bnd = (j == P.ny - 1) && !P.per_y;
                if (nj >= P.ny) nj -= P.ny;
            }
            // face geometry of the own copy and of the neighbour's copy
            double m0[3], m1[3];
            auto face_m = [&](int fi, int fj, int f, double* m)
            {
                if (f < 2)
                {
                    if (P.curved)
                    {
                        const double* fm =
                            P.fxm + ((static_cast<long long>(fj) * (P.nx + 1) + fi + f) * N + idx) * 4;
                        m[0] = fm[0]; m[1] = fm[1]; m[2] = fm[2];
                    }
                    else
                    {
                        m[0] = 1.0; m[1] = 0.0; m[2] = P.sc_x;
                    }
                }
                else
                {
                    if (P.curved)
                    {
                        const double* fm =
                            P.fym + ((static_cast<long long>(fj + (f - 2)) * P.nx + fi) * N + idx) * 4;
                        m[0] = fm[0]; m[1] = fm[1]; m[2] = fm[2];
                    }
                    else
                    {
                        m[0] = 0.0; m[1] = 1.0; m[2] = P.sc_y;
                    }
                }
            };
            face_m(i, j, face, m0);
            FaceAdjIn in;
            const long long own = (eg * NN + ol * N + ok) * 4;
            ld4d(u + own, in.uo);
            in.nx        = m0[0];
            in.ny        = m0[1];
            in.own_right = (face == 0 || face == 2);
            in.bnd       = bnd;
            in.dir       = dir;
            const double* wbo = wb + ((eg * 4 + face) * N + idx) * 4;
            double        cs[4], cn[4], ub[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                cs[q] = m0[2] * wbo[q];
                cn[q] = -m0[2] * wbo[q];
            }
            if (!bnd)
            {
                const long long en = static_cast<long long>(nj) * P.nx + ni;
                ld4d(u + (en * NN + nl * N + nk) * 4, in.un);
                face_m(ni, nj, nface, m1);
                const double* wbn = wb + ((en * 4 + nface) * N + idx) * 4;
                if (m1[0] == m0[0] && m1[1] == m0[1] && m1[2] == m0[2])
                {
#pragma unroll
                    for (int q = 0; q < 4; ++q) cs[q] += m0[2] * wbn[q];
                }
                else
                {
                    // periodic wrap on a map whose two ends differ: the neighbour's
                    // copy has its own normal and scale
                    FaceAdjIn in2 = in;
                    in2.nx        = m1[0];
                    in2.ny        = m1[1];
                    double c2[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) c2[q] = m1[2] * wbn[q];
                    riemann_adjoint(P.gas, P.riemann, in2, c2, ub);
                }
            }
            face_adjoint(P.gas, P.riemann, P.bc, in, cs, cn, ub);
#pragma unroll
            for (int q = 0; q < 4; ++q) sFc[fe][face][idx][q] = ub[q];
        }
        __syncthreads();
        if (!act) return;
        // ---- volume: Fbar, Gbar at this node ----
        double Fb[4], Gb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            double f = 0.0, g = 0.0;
            for (int m = 0; m < N; ++m)
            {
                // d/dxi terms: nodes (m, l) with D[m][k]; d/deta terms: nodes (k, m) with D[m][l]
                f += sD[m * N + k] * sA[el][l * N + m][q] + sD[m * N + l] * sB[el][m * N + k][q];
                g += sD[m * N + k] * sC[el][l * N + m][q] + sD[m * N + l] * sDd[el][m * N + k][q];
            }
            Fb[q] = f;
            Gb[q] = g;
        }
        double ub[4] = {0, 0, 0, 0};
        flux2_vjp(P.gas, U, Fb, Gb, ub);
        if (P.rinv)
        {
            // radial source (pointwise): one forward-dual pass per state component
            const double ir = P.rinv[e * NN + node];
            for (int i = 0; i < 4; ++i)
            {
                Dual A[4], sr[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) A[q] = {U[q], q == i ? 1.0 : 0.0};
                radial_source2(P.gas, A, ir, sr);
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc += sr[q].d * W[q];
                ub[i] += acc;
            }
        }
        // ---- gather face contributions of this node ----
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            double s = ub[q];
            if (k == 0) s += sFc[el][0][l][q];
            if (k == N - 1) s += sFc[el][1][l][q];
            if (l == 0) s += sFc[el][2][k][q];
            if (l == N - 1) s += sFc[el][3][k][q];
            ub[q] = ao.at((e * NN + node) * 4 + q, W[q], s);
        }
        st4d(out + (e * NN + node) * 4, ub);
    }

    // ------------------------------------------------------------------
    // 1D adjoint, one thread per element (physics_impl_1D.hpp:123-181 transposed)
    // ------------------------------------------------------------------
    __device__ __forceinline__ void riemann1_adj(const GasP& g, int kind, const double a[3], const double b[3],
                                                 int wrt_b, const double c[3], double ub[3], const BCP* bc,
                                                 int side)
    {
        // ub += (d F*(a, b) / d own)^T c ; own = b if wrt_b else a.  If bc, the other state is
        // bc_exterior1(own) (chained).
        for (int i = 0; i < 3; ++i)
        {
            Dual A[3], B[3], Fs[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                A[q] = {a[q], (!wrt_b && q == i) ? 1.0 : 0.0};
                B[q] = {b[q], (wrt_b && q == i) ? 1.0 : 0.0};
            }
            if (bc)
            {
                if (wrt_b)
                    bc_exterior1(*bc, side, B, A);
                else
                    bc_exterior1(*bc, side, A, B);
            }
            riemann1(g, kind, A, B, Fs);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) s += Fs[q].d * c[q];
            ub[i] += s;
        }
    }
    __device__ __forceinline__ void flux1_adj(const GasP& g, const double a[3], const double c[3], double ub[3])
    {
        for (int i = 0; i < 3; ++i)
        {
            Dual A[3], F[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) A[q] = {a[q], q == i ? 1.0 : 0.0};
            flux1(g, A, F);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) s += F[q].d * c[q];
            ub[i] += s;
        }
    }
    template <int N>
    __global__ __launch_bounds__(64) void k_adj1d(const Res1DP P, const double* __restrict__ u,
                                                   const double* __restrict__ w, const AdjOut ao,
                                                   double* __restrict__ out)
    {
        const int e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= P.ne) return;
        const double* D  = P.tab;
        const double* dL = P.tab + N * N;
        const double* dR = P.tab + N * N + N;
        const long long b = static_cast<long long>(e) * 3 * N;
        const int ep = (e == 0) ? P.ne - 1 : e - 1, en = (e == P.ne - 1) ? 0 : e + 1;
        const bool lb = (e == 0) && !P.periodic, rb = (e == P.ne - 1) && !P.periodic;
        double ub[3][N];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) ub[q][m] = 0.0;
        // volume: R_i += ij * D[i][j] F_j
        for (int jn = 0; jn < N; ++jn)
        {
            double c[3], a[3], t[3] = {0, 0, 0};
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                double s = 0.0;
                for (int in = 0; in < N; ++in) s += D[in * N + jn] * w[b + q * N + in];
                c[q] = P.inv_jac * s;
                a[q] = u[b + q * N + jn];
            }
            flux1_adj(P.gas, a, c, t);
#pragma unroll
            for (int q = 0; q < 3; ++q) ub[q][jn] += t[q];
        }
        // lifted cotangents of own faces and neighbours' copies
        double cL[3], cR[3], cLn[3], cRp[3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
        {
            double sl = 0, sr = 0, sln = 0, srp = 0;
            for (int in = 0; in < N; ++in)
            {
                sl += dL[in] * w[b + q * N + in];
                sr += dR[in] * w[b + q * N + in];
                sln += dL[in] * w[static_cast<long long>(en) * 3 * N + q * N + in];
                srp += dR[in] * w[static_cast<long long>(ep) * 3 * N + q * N + in];
            }
            cL[q]  = P.inv_jac * sl;
            cR[q]  = P.inv_jac * sr;
            cLn[q] = P.inv_jac * sln;
            cRp[q] = P.inv_jac * srp;
        }
        double first[3], last[3], prev[3], next[3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
        {
            first[q] = u[b + q * N];
            last[q]  = u[b + q * N + N - 1];
            prev[q]  = u[static_cast<long long>(ep) * 3 * N + q * N + N - 1];
            next[q]  = u[static_cast<long long>(en) * 3 * N + q * N];
        }
        // left face: F*(prev, first) appears in R_e (cL) and, if interior, in R_{e-1} (cRp)
        {
            double c[3], t[3] = {0, 0, 0}, mc[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                c[q]  = cL[q] + (lb ? 0.0 : cRp[q]);
                mc[q] = -cL[q];
            }
            riemann1_adj(P.gas, P.riemann, prev, first, 1, c, t, lb ? &P.bc : nullptr, 0);
            flux1_adj(P.gas, first, mc, t);
#pragma unroll
            for (int q = 0; q < 3; ++q) ub[q][0] += t[q];
        }
        // right face: F*(last, next) appears in R_e (cR) and, if interior, in R_{e+1} (cLn)
        {
            double c[3], t[3] = {0, 0, 0}, mc[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                c[q]  = cR[q] + (rb ? 0.0 : cLn[q]);
                mc[q] = -cR[q];
            }
            riemann1_adj(P.gas, P.riemann, last, next, 0, c, t, rb ? &P.bc : nullptr, 1);
            flux1_adj(P.gas, last, mc, t);
#pragma unroll
            for (int q = 0; q < 3; ++q) ub[q][N - 1] += t[q];
        }
        if (P.rinv && (P.periodic || (e > 0 && e < P.ne - 1))) // interior elements only, as the residual
        {
            for (int m = 0; m < N; ++m)
            {
                const double ir = P.rinv[static_cast<long long>(e) * N + m];
                for (int i = 0; i < 3; ++i)
                {
                    Dual A[3], sr[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) A[q] = {u[b + q * N + m], q == i ? 1.0 : 0.0};
                    radial_source1(P.gas, A, ir, sr);
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc += sr[q].d * w[b + q * N + m];
                    ub[i][m] += acc;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) out[b + q * N + m] = ao.at(b + q * N + m, w[b + q * N + m], ub[q][m]);
    }
} // namespace frsdg

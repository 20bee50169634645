// frsdg_kernels.hpp — HIP kernels of the FR residual path for gfx950.
//
// Residual kernel design (2D, one launch per residual):
//   * a workgroup owns EPB consecutive elements (EPB = floor(BS / n^2)),
//     one thread per solution node; the element blocks are contiguous in HBM
//     so the 32-byte node loads/stores of a workgroup are fully coalesced;
//   * each thread evaluates the Euler fluxes F,G of its node once and stages
//     them in LDS; the derivative-matrix contraction reads them from LDS
//     (D itself is staged in LDS too);
//   * "owner computes" faces: every element evaluates the Riemann flux of
//     its own four faces (4n face-node tasks), gathering the neighbour trace
//     straight from global memory.  A shared face is evaluated by both
//     neighbours with identical arguments, so the result is deterministic
//     and conservative without atomics;
//   * the VCJH lift and the volume term are combined per node and written
//     once; optional output policies fuse rhs = -R, F = udot + R and the
//     SSP-RK stage update into the same store;
//   * affine meshes use the four constant metric terms (only the
//     d/dxi of F and d/deta of G survive, as in the reference where the
//     cross terms are exact zeros); curved meshes read per-node metrics and
//     per-face (nx, ny, scale) precomputed once in HBM;
//   * block index -> element tile is remapped so that each XCD walks a
//     contiguous band of elements (neighbour traces hit its own L2).
// The same kernel body is instantiated on double (residual), Dual (forward
// tangent J*v) and Cplx (complex step), see frsdg_scalar.hpp.
#pragma once
#include "frsdg_geometry.hpp"
#include "frsdg_physics.hpp"

namespace frsdg
{
    struct Res2DP
    {
        int       nx, ny;
        long long E;
        int       per_x, per_y;
        int       curved;
        int       riemann;
        double    xi_x, eta_y, sc_x, sc_y; // affine constants
        const double* __restrict__ vmet;   // [E][n*n][4]: xi_x, xi_y, eta_x, eta_y
        const double* __restrict__ fxm;    // [ny][nx+1][n][4]: nx, ny, scale, 0
        const double* __restrict__ fym;    // [ny+1][nx][n][4]
        const double* __restrict__ tab;    // D[n*n], dL[n], dR[n], filt[n*n]
        GasP      gas;
        BCP       bc;
    };

    // blockIdx -> logical tile so that XCD x (= b % 8) processes one contiguous band
    __device__ __forceinline__ long long xcd_tile(long long b, long long G)
    {
        const long long x = b & 7, q = G >> 3, r = G & 7;
        return x * q + (x < r ? x : r) + (b >> 3);
    }

    // ------------------------- I/O policies -------------------------
    __device__ __forceinline__ void ld4d(const double* __restrict__ p, double o[4])
    {
        const double2 a = reinterpret_cast<const double2*>(p)[0];
        const double2 b = reinterpret_cast<const double2*>(p)[1];
        o[0] = a.x;
        o[1] = a.y;
        o[2] = b.x;
        o[3] = b.y;
    }
    __device__ __forceinline__ void st4d(double* __restrict__ p, const double o[4])
    {
        reinterpret_cast<double2*>(p)[0] = make_double2(o[0], o[1]);
        reinterpret_cast<double2*>(p)[1] = make_double2(o[2], o[3]);
    }

    // out = add + scale * R   (residual: scale 1; rhs: scale -1; ifunction: add = udot)
    struct IORes
    {
        using S = double;
        static constexpr bool can_filter = false;
        const double* __restrict__ u;
        double* __restrict__ out;
        const double* __restrict__ add;
        double scale;
        __device__ double ld(long long i) const { return u[i]; }
        __device__ void ld4(long long i, double o[4]) const { ld4d(u + i, o); }
        __device__ void st(long long i, double, double r) const
        {
            out[i] = add ? add[i] + r : scale * r;
        }
        __device__ void st4(long long i, const double*, const double r[4]) const
        {
            double o[4];
            if (add)
            {
                double a[4];
                ld4d(add + i, a);
                for (int q = 0; q < 4; ++q) o[q] = a[q] + r[q];
            }
            else
                for (int q = 0; q < 4; ++q) o[q] = scale * r[q];
            st4d(out + i, o);
        }
    };
    // SSP-RK stage: out = a0*u0 + b*(us + cdt*(-R(us)))  (u0 read only if a0 != 0)
    struct IOStage
    {
        using S = double;
        static constexpr bool can_filter = true;
        const double* __restrict__ u;  // us
        const double* u0; // may alias out (each node read by its own thread before any write)
        double* out;
        double  a0, b, cdt;
        __device__ double ld(long long i) const { return u[i]; }
        __device__ void ld4(long long i, double o[4]) const { ld4d(u + i, o); }
        __device__ void st(long long i, double own, double r) const
        {
            const double s = own + cdt * (-r);
            out[i]         = (a0 != 0.0) ? a0 * u0[i] + b * s : b * s;
        }
        __device__ void val4(long long i, const double own[4], const double r[4], double o[4]) const
        {
            for (int q = 0; q < 4; ++q) o[q] = own[q] + cdt * (-r[q]);
            if (a0 != 0.0)
            {
                double z[4];
                ld4d(u0 + i, z);
                for (int q = 0; q < 4; ++q) o[q] = a0 * z[q] + b * o[q];
            }
            else
                for (int q = 0; q < 4; ++q) o[q] = b * o[q];
        }
        __device__ void put4(long long i, const double o[4]) const { st4d(out + i, o); }
        __device__ void st4(long long i, const double own[4], const double r[4]) const
        {
            double o[4];
            val4(i, own, r, o);
            st4d(out + i, o);
        }
    };
    // forward tangent: out = sigma*v + dR
    struct IOTan
    {
        using S = Dual;
        static constexpr bool can_filter = false;
        const double* __restrict__ u;
        const double* __restrict__ v;
        double* __restrict__ out;
        double sigma;
        __device__ Dual ld(long long i) const { return {u[i], v[i]}; }
        __device__ void ld4(long long i, Dual o[4]) const
        {
            double a[4], b[4];
            ld4d(u + i, a);
            ld4d(v + i, b);
            for (int q = 0; q < 4; ++q) o[q] = {a[q], b[q]};
        }
        __device__ void st(long long i, Dual own, Dual r) const { out[i] = sigma * own.d + r.d; }
        __device__ void st4(long long i, const Dual own[4], const Dual r[4]) const
        {
            double o[4];
            for (int q = 0; q < 4; ++q) o[q] = sigma * own[q].d + r[q].d;
            st4d(out + i, o);
        }
    };
    // complex step (petsc_solver_impl.hpp:289-356): u + i h v; out = sigma Im(u_iv)/h + Im R/h
    struct IOCs
    {
        using S = Cplx;
        static constexpr bool can_filter = false;
        const double* __restrict__ u;
        const double* __restrict__ v;
        double* __restrict__ out;
        double sigma, h, inv_h;
        __device__ Cplx ld(long long i) const { return {u[i], v[i] * h}; }
        __device__ void ld4(long long i, Cplx o[4]) const
        {
            double a[4], b[4];
            ld4d(u + i, a);
            ld4d(v + i, b);
            for (int q = 0; q < 4; ++q) o[q] = {a[q], b[q] * h};
        }
        __device__ void st(long long i, Cplx own, Cplx r) const
        {
            out[i] = sigma * own.im * inv_h + r.im * inv_h;
        }
        __device__ void st4(long long i, const Cplx own[4], const Cplx r[4]) const
        {
            double o[4];
            for (int q = 0; q < 4; ++q) o[q] = sigma * own[q].im * inv_h + r[q].im * inv_h;
            st4d(out + i, o);
        }
    };

    template <class S>
    struct BlockSize
    {
        static constexpr int value = 256;
    };
    template <>
    struct BlockSize<Dual>
    {
        static constexpr int value = 128;
    };
    template <>
    struct BlockSize<Cplx>
    {
        static constexpr int value = 128;
    };
    template <int N, class S>
    struct Tile2D
    {
        static constexpr int BS  = BlockSize<S>::value;
        static constexpr int NN  = N * N;
        static constexpr int EPB = (BS / NN) > 0 ? (BS / NN) : 1;
    };

    // ------------------------------------------------------------------
    // 2D residual: physics_impl_2D.hpp:267-357 (volume) + :689-899 (faces)
    // ------------------------------------------------------------------
    // LDS of one residual tile
    template <int N, class S>
    struct TileSmem
    {
        static constexpr int NN = N * N, EPB = Tile2D<N, S>::EPB;
        S      U[EPB][NN][4];
        S      F[EPB][NN][4];
        S      G[EPB][NN][4];
        S      Del[EPB][4][N][4];
        double D[NN], dL[N], dR[N], Fm[NN];
        int    I[EPB], J[EPB]; // element coordinates of the tile
    };

    // One tile = EPB consecutive elements of one mesh, processed by one
    // workgroup.  `epb` <= EPB elements are active (the batched sweep may run
    // smaller tiles).  With `filt` (stage outputs only) the element-local
    // spectral filter of the post-step hook is applied in LDS before the store.
    template <int N, class IO>
    __device__ __forceinline__ void res2d_tile(const Res2DP& P, const IO& io, long long blk, int epb, bool filt,
                                               TileSmem<N, typename IO::S>& sm)
    {
        using S           = typename IO::S;
        constexpr int BS  = Tile2D<N, S>::BS;
        constexpr int NN  = N * N;
        const int     tid = threadIdx.x;
        for (int q = tid; q < NN; q += BS)
        {
            sm.D[q] = P.tab[q];
            if (filt) sm.Fm[q] = P.tab[NN + 2 * N + q];
        }
        if (tid < N)
        {
            sm.dL[tid] = P.tab[NN + tid];
            sm.dR[tid] = P.tab[NN + N + tid];
        }
        const int       el   = tid / NN;
        const int       node = tid - el * NN;
        const int       k    = node % N;
        const int       l    = node / N;
        const long long e    = blk * epb + el;
        const bool      act  = (el < epb) && (e < P.E);
        if (tid < epb)
        {
            // element coordinates once per element (32-bit: E < 2^31 is checked at setup)
            const int eg = static_cast<int>(blk * epb + tid);
            sm.J[tid]    = eg / P.nx;
            sm.I[tid]    = eg - sm.J[tid] * P.nx;
        }

        S own[4];
        if (act)
        {
            io.ld4((e * NN + node) * 4, own);
            S F[4], G[4];
            flux2(P.gas, own, F, G);
            for (int q = 0; q < 4; ++q)
            {
                sm.U[el][node][q] = own[q];
                sm.F[el][node][q] = F[q];
                sm.G[el][node][q] = G[q];
            }
        }
        __syncthreads();

        // ---- face tasks: 4 faces x N nodes per element ----
        for (int t = tid; t < epb * 4 * N; t += BS)
        {
            const int       fe = t / (4 * N);
            const int       ft = t - fe * 4 * N;
            const long long eg = blk * epb + fe;
            if (eg >= P.E) continue;
            const int face = ft / N, idx = ft - (ft / N) * N;
            const int i = sm.I[fe], j = sm.J[fe];
            int       ok, ol, nk, nl, ni = i, nj = j, dir;
            bool      bnd;
            if (face == 0)
            {
                ok = 0; ol = idx; nk = N - 1; nl = idx; ni = i - 1; dir = 0;
                bnd = (i == 0) && !P.per_x;
                if (ni < 0) ni += P.nx;
            }
            else if (face == 1)
            {
                ok = N - 1; ol = idx; nk = 0; nl = idx; ni = i + 1; dir = 1;
                bnd = (i == P.nx - 1) && !P.per_x;
                if (ni >= P.nx) ni -= P.nx;
            }
            else if (face == 2)
            {
                ok = idx; ol = 0; nk = idx; nl = N - 1; nj = j - 1; dir = 3;
                bnd = (j == 0) && !P.per_y;
                if (nj < 0) nj += P.ny;
            }
            else
            {
                ok = idx; ol = N - 1; nk = idx; nl = 0; nj = j + 1; dir = 2;
                bnd = (j == P.ny - 1) && !P.per_y;
                if (nj >= P.ny) nj -= P.ny;
            }
            const int on = ol * N + ok;
            double    fnx, fny, sc;
            if (face < 2)
            {
                if (P.curved)
                {
                    const double* fm =
                        P.fxm + ((static_cast<long long>(j) * (P.nx + 1) + i + face) * N + idx) * 4;
                    fnx = fm[0];
                    fny = fm[1];
                    sc  = fm[2];
                }
                else
                {
                    fnx = 1.0;
                    fny = 0.0;
                    sc  = P.sc_x;
                }
            }
            else
            {
                if (P.curved)
                {
                    const double* fm =
                        P.fym + ((static_cast<long long>(j + (face - 2)) * P.nx + i) * N + idx) * 4;
                    fnx = fm[0];
                    fny = fm[1];
                    sc  = fm[2];
                }
                else
                {
                    fnx = 0.0;
                    fny = 1.0;
                    sc  = P.sc_y;
                }
            }
            S Uo[4], Un[4], Fs[4];
            for (int q = 0; q < 4; ++q) Uo[q] = sm.U[fe][on][q];
            if (bnd)
                bc_exterior(P.bc, dir, Uo, fnx, fny, Un);
            else
                io.ld4(((static_cast<long long>(nj) * P.nx + ni) * NN + nl * N + nk) * 4, Un);
            if (face == 0 || face == 2)
                riemann2(P.gas, P.riemann, Un, Uo, fnx, fny, Fs);
            else
                riemann2(P.gas, P.riemann, Uo, Un, fnx, fny, Fs);
            for (int q = 0; q < 4; ++q)
            {
                const S fn               = sm.F[fe][on][q] * fnx + sm.G[fe][on][q] * fny;
                sm.Del[fe][face][idx][q] = sc * (Fs[q] - fn);
            }
        }
        __syncthreads();

        S r[4];
        if (act)
        {
            if (!P.curved)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    S dFx = Sc<S>::c(0.0), dGy = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m)
                    {
                        dFx = dFx + sm.D[k * N + m] * sm.F[el][l * N + m][q];
                        dGy = dGy + sm.D[l * N + m] * sm.G[el][m * N + k][q];
                    }
                    r[q] = P.xi_x * dFx + P.eta_y * dGy;
                }
            }
            else
            {
                double mt[4];
                ld4d(P.vmet + (e * NN + node) * 4, mt);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    S dFx = Sc<S>::c(0.0), dGx = Sc<S>::c(0.0), dFy = Sc<S>::c(0.0), dGy = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m)
                    {
                        const double a = sm.D[k * N + m], b = sm.D[l * N + m];
                        dGx = dGx + a * sm.G[el][l * N + m][q];
                        dFx = dFx + a * sm.F[el][l * N + m][q];
                        dGy = dGy + b * sm.G[el][m * N + k][q];
                        dFy = dFy + b * sm.F[el][m * N + k][q];
                    }
                    const S px = mt[0] * dFx + mt[2] * dFy;
                    const S py = mt[1] * dGx + mt[3] * dGy;
                    r[q]       = px + py;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                r[q] = r[q] + sm.Del[el][0][l][q] * sm.dL[k];
                r[q] = r[q] + sm.Del[el][1][l][q] * sm.dR[k];
                r[q] = r[q] + sm.Del[el][2][k][q] * sm.dL[l];
                r[q] = r[q] + sm.Del[el][3][k][q] * sm.dR[l];
            }
        }
        if constexpr (IO::can_filter)
        {
            if (filt)
            {
                // stage value -> LDS (U is free: its last readers, the face
                // tasks, finished at the previous barrier), then
                // out = 1/2 (F_x v + F_y v)
                if (act)
                {
                    double o[4];
                    io.val4((e * NN + node) * 4, own, r, o);
                    for (int q = 0; q < 4; ++q) sm.U[el][node][q] = o[q];
                }
                __syncthreads();
                if (act)
                {
                    double o[4];
                    for (int q = 0; q < 4; ++q)
                    {
                        double sx = 0.0, sy = 0.0;
                        for (int c = 0; c < N; ++c)
                        {
                            sx += sm.Fm[k * N + c] * sm.U[el][l * N + c][q];
                            sy += sm.Fm[l * N + c] * sm.U[el][c * N + k][q];
                        }
                        o[q] = 0.5 * (sx + sy);
                    }
                    io.put4((e * NN + node) * 4, o);
                }
                return;
            }
        }
        if (act) io.st4((e * NN + node) * 4, own, r);
    }

    // ------------------------------------------------------------------
    // 2D residual: physics_impl_2D.hpp:267-357 (volume) + :689-899 (faces)
    // ------------------------------------------------------------------
    template <int N, class IO>
    __global__ __launch_bounds__(BlockSize<typename IO::S>::value) void k_res2d(const Res2DP P, const IO io)
    {
        __shared__ TileSmem<N, typename IO::S> sm;
        res2d_tile<N, IO>(P, io, xcd_tile(blockIdx.x, gridDim.x), Tile2D<N, typename IO::S>::EPB, false, sm);
    }
    // last SSP-RK stage fused with the post-step spectral filter
    template <int N>
    __global__ __launch_bounds__(256) void k_res2d_stage_filter(const Res2DP P, const IOStage io)
    {
        __shared__ TileSmem<N, double> sm;
        res2d_tile<N, IOStage>(P, io, xcd_tile(blockIdx.x, gridDim.x), Tile2D<N, double>::EPB, true, sm);
    }

    // ------------------------------------------------------------------
    // 1D residual: physics_impl_1D.hpp:8-94 + local_steady_res :123-181.
    // One thread per element (the 1D configurations are launch-bound).
    // ------------------------------------------------------------------
    struct Res1DP
    {
        int    ne;
        int    periodic;
        int    riemann;
        double inv_jac;
        const double* __restrict__ tab;
        GasP   gas;
        BCP    bc;
    };
    template <int N, class IO>
    __global__ __launch_bounds__(64) void k_res1d(const Res1DP P, const IO io)
    {
        using S        = typename IO::S;
        const int    e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= P.ne) return;
        const double* D  = P.tab;
        const double* dL = P.tab + N * N;
        const double* dR = P.tab + N * N + N;
        const long long base = static_cast<long long>(e) * 3 * N;
        S u[3][N];
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) u[q][m] = io.ld(base + q * N + m);
        S prev[3], next[3];
        const int ep = (e == 0) ? P.ne - 1 : e - 1;
        const int en = (e == P.ne - 1) ? 0 : e + 1;
        S         first[3] = {u[0][0], u[1][0], u[2][0]};
        S         last[3]  = {u[0][N - 1], u[1][N - 1], u[2][N - 1]};
        if (e == 0 && !P.periodic)
            bc_exterior1(P.bc, 0, first, prev);
        else
            for (int q = 0; q < 3; ++q) prev[q] = io.ld(static_cast<long long>(ep) * 3 * N + q * N + (N - 1));
        if (e == P.ne - 1 && !P.periodic)
            bc_exterior1(P.bc, 1, last, next);
        else
            for (int q = 0; q < 3; ++q) next[q] = io.ld(static_cast<long long>(en) * 3 * N + q * N);
        S res[3][N];
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) res[q][m] = Sc<S>::c(0.0);
        for (int jn = 0; jn < N; ++jn)
        {
            S nd[3] = {u[0][jn], u[1][jn], u[2][jn]}, F[3];
            flux1(P.gas, nd, F);
            for (int in = 0; in < N; ++in)
            {
                const double Dij = P.inv_jac * D[in * N + jn];
                for (int q = 0; q < 3; ++q) res[q][in] = res[q][in] + Dij * F[q];
            }
        }
        S FdL[3], FdR[3], FsL[3], FsR[3];
        flux1(P.gas, first, FdL);
        flux1(P.gas, last, FdR);
        riemann1(P.gas, P.riemann, prev, first, FsL);
        riemann1(P.gas, P.riemann, last, next, FsR);
        for (int q = 0; q < 3; ++q)
        {
            const S dl = FsL[q] - FdL[q], dr = FsR[q] - FdR[q];
            for (int in = 0; in < N; ++in)
                res[q][in] = res[q][in] + P.inv_jac * (dl * dL[in] + dr * dR[in]);
        }
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) io.st(base + q * N + m, u[q][m], res[q][m]);
    }

    // ------------------------------------------------------------------
    // Spectral filter, limit_solution semantics (physics_impl_2D.hpp:76-124):
    // out = 1/2 (F_x u + F_y u) per equation; element-local, so in place is safe.
    // ------------------------------------------------------------------
    template <int N>
    __global__ __launch_bounds__(256) void k_filter2d(long long E, const double* __restrict__ tab,
                                                       const double* u, double* out, int transpose)
    {
        constexpr int NN = N * N, EPB = (256 / NN) > 0 ? 256 / NN : 1;
        __shared__ double sU[EPB][NN][4];
        __shared__ double sFm[NN];
        const int tid = threadIdx.x;
        for (int q = tid; q < NN; q += 256)
        {
            const int r = q / N, c = q - (q / N) * N; // transposed filter for the adjoint
            sFm[q] = tab[NN + 2 * N + (transpose ? c * N + r : q)];
        }
        const long long blk = xcd_tile(blockIdx.x, gridDim.x);
        const int el = tid / NN, node = tid - el * NN, k = node % N, l = node / N;
        const long long e   = blk * EPB + el;
        const bool      act = (el < EPB) && (e < E);
        if (act)
        {
            double o[4];
            ld4d(u + (e * NN + node) * 4, o);
            for (int q = 0; q < 4; ++q) sU[el][node][q] = o[q];
        }
        __syncthreads();
        if (!act) return;
        double o[4];
        for (int q = 0; q < 4; ++q)
        {
            double sx = 0.0, sy = 0.0;
            for (int c = 0; c < N; ++c)
            {
                sx += sFm[k * N + c] * sU[el][l * N + c][q];
                sy += sFm[l * N + c] * sU[el][c * N + k][q];
            }
            o[q] = 0.5 * (sx + sy);
        }
        st4d(out + (e * NN + node) * 4, o);
    }
    template <int N>
    __global__ __launch_bounds__(64) void k_filter1d(int ne, const double* __restrict__ tab,
                                                      const double* u, double* out, int transpose)
    {
        const int e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= ne) return;
        const double* Fm = tab + N * N + 2 * N;
        double        x[3 * N];
        for (int q = 0; q < 3 * N; ++q) x[q] = u[static_cast<long long>(e) * 3 * N + q];
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < N; ++r)
            {
                double s = 0.0;
                for (int c = 0; c < N; ++c) s += Fm[transpose ? c * N + r : r * N + c] * x[q * N + c];
                out[static_cast<long long>(e) * 3 * N + q * N + r] = s;
            }
    }
} // namespace frsdg

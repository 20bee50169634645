// frsdg_kernels.hpp — HIP kernels of the FR residual path for gfx950.
//
// Residual kernel design (2D, one launch per residual):
//   * a workgroup owns EPB consecutive elements (EPB = floor(BS / n^2)),
//     one thread per solution node; the element blocks are contiguous in HBM
//     so the 32-byte node loads/stores of a workgroup are fully coalesced;
//   * each thread evaluates the Euler fluxes F,G of its node once and stages
//     them in LDS; the derivative-matrix contraction then runs as line tasks
//     (one element row or column, one equation per thread): the line is read
//     from LDS once into registers, D comes through the scalar cache, and the
//     metric-weighted derivatives of all N nodes of the line go back to LDS,
//     so each flux value is read twice instead of 2N (4N curved) times;
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
        const double* __restrict__ vmet;   // [E][4][n*n] (vmet_node)
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
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = a[q] + r[q];
            }
            else
#pragma unroll
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
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = own[q] + cdt * (-r[q]);
            if (a0 != 0.0)
            {
                double z[4];
                ld4d(u0 + i, z);
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = a0 * z[q] + b * o[q];
            }
            else
#pragma unroll
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
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = {a[q], b[q]};
        }
        __device__ void st(long long i, Dual own, Dual r) const { out[i] = sigma * own.d + r.d; }
        __device__ void st4(long long i, const Dual own[4], const Dual r[4]) const
        {
            double o[4];
#pragma unroll
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
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = {a[q], b[q] * h};
        }
        __device__ void st(long long i, Cplx own, Cplx r) const
        {
            out[i] = sigma * own.im * inv_h + r.im * inv_h;
        }
        __device__ void st4(long long i, const Cplx own[4], const Cplx r[4]) const
        {
            double o[4];
#pragma unroll
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
    // LDS of one residual tile.  Node (k, l) of tile element el, equation q,
    // lives at X[el][l][k * 4 + q]; rows are padded so that the row stride is
    // an odd multiple of 32 bytes (line tasks of one wavefront then spread
    // over all LDS banks).
    template <int N, class S>
    struct TileSmem
    {
        static constexpr int NN = N * N, EPB = Tile2D<N, S>::EPB;
        static constexpr int RB = N * 4 * static_cast<int>(sizeof(S));                 // row bytes
        static constexpr int RP = N * 4 + ((RB % 64) == 0 ? 32 / static_cast<int>(sizeof(S)) : 0);
        S      U[EPB][N][RP]; // state; after the face phase: xi-line part of the volume term
        S      F[EPB][N][RP];
        S      G[EPB][N][RP];
        S      V[EPB][N][RP]; // eta-line part of the volume term
        S      Del[EPB][4][N][4];
        double dL[N], dR[N], Fm[NN];
    };

    // Metrics of node (k, l) of element e.  Layout [E][4][NN]: xi_x, xi_y
    // row-major (l * N + k), eta_x, eta_y column-major (k * N + l), so that a
    // xi-line reads its row and an eta-line its column contiguously.
    template <int N>
    __device__ __forceinline__ void vmet_node(const double* __restrict__ vm, long long e, int k, int l, double mt[4])
    {
        constexpr int NN = N * N;
        const double* b  = vm + e * 4 * NN;
        mt[0]            = b[l * N + k];
        mt[1]            = b[NN + l * N + k];
        mt[2]            = b[2 * NN + k * N + l];
        mt[3]            = b[3 * NN + k * N + l];
    }

    // One face-node task of a tile (element fe, face, node idx along the face):
    // geometry, and for interior faces the neighbour's trace gathered from HBM.
    template <int N, class S>
    struct FaceTask
    {
        int    fe, face, idx, ok, ol, dir;
        bool   live, bnd;
        double fnx, fny, sc;
        S      Un[4];
    };

    // element coordinates of the first element of a tile (uniform; 32-bit:
    // E < 2^31 is checked at setup)
    struct TileOrigin
    {
        int i0, j0;
    };
    __device__ __forceinline__ TileOrigin tile_origin(const Res2DP& P, long long blk, int epb)
    {
        const int e0 = static_cast<int>(blk * epb);
        const int j0 = e0 / P.nx;
        return {e0 - j0 * P.nx, j0};
    }

    template <int N, class IO>
    __device__ __forceinline__ void face_task(const Res2DP& P, const IO& io, long long blk, int epb, TileOrigin o,
                                              int t, FaceTask<N, typename IO::S>& f)
    {
        constexpr int NN   = N * N;
        f.fe               = t / (4 * N);
        const int       ft = t - f.fe * 4 * N;
        const long long eg = blk * epb + f.fe;
        f.live             = (f.fe < epb) && (eg < P.E);
        if (!f.live) return;
        f.face = ft / N;
        f.idx  = ft - f.face * N;
        // element coordinates: walk from the tile origin (one wrap at most
        // when nx >= epb)
        int i = o.i0 + f.fe, j = o.j0;
        while (i >= P.nx)
        {
            i -= P.nx;
            ++j;
        }
        const int idx = f.idx;
        int       nk, nl, ni = i, nj = j;
        if (f.face == 0)
        {
            f.ok = 0; f.ol = idx; nk = N - 1; nl = idx; ni = i - 1; f.dir = 0;
            f.bnd = (i == 0) && !P.per_x;
            if (ni < 0) ni += P.nx;
        }
        else if (f.face == 1)
        {
            f.ok = N - 1; f.ol = idx; nk = 0; nl = idx; ni = i + 1; f.dir = 1;
            f.bnd = (i == P.nx - 1) && !P.per_x;
            if (ni >= P.nx) ni -= P.nx;
        }
        else if (f.face == 2)
        {
            f.ok = idx; f.ol = 0; nk = idx; nl = N - 1; nj = j - 1; f.dir = 3;
            f.bnd = (j == 0) && !P.per_y;
            if (nj < 0) nj += P.ny;
        }
        else
        {
            f.ok = idx; f.ol = N - 1; nk = idx; nl = 0; nj = j + 1; f.dir = 2;
            f.bnd = (j == P.ny - 1) && !P.per_y;
            if (nj >= P.ny) nj -= P.ny;
        }
        if (f.face < 2)
        {
            if (P.curved)
            {
                const double* fm = P.fxm + ((static_cast<long long>(j) * (P.nx + 1) + i + f.face) * N + idx) * 4;
                f.fnx = fm[0];
                f.fny = fm[1];
                f.sc  = fm[2];
            }
            else
            {
                f.fnx = 1.0;
                f.fny = 0.0;
                f.sc  = P.sc_x;
            }
        }
        else
        {
            if (P.curved)
            {
                const double* fm = P.fym + ((static_cast<long long>(j + (f.face - 2)) * P.nx + i) * N + idx) * 4;
                f.fnx = fm[0];
                f.fny = fm[1];
                f.sc  = fm[2];
            }
            else
            {
                f.fnx = 0.0;
                f.fny = 1.0;
                f.sc  = P.sc_y;
            }
        }
        if (!f.bnd) io.ld4(((static_cast<long long>(nj) * P.nx + ni) * NN + nl * N + nk) * 4, f.Un);
    }

    // Riemann flux of a face task -> scaled flux jump in sm.Del
    template <int N, class IO>
    __device__ __forceinline__ void face_flux(const Res2DP& P, FaceTask<N, typename IO::S>& f,
                                              TileSmem<N, typename IO::S>& sm)
    {
        using S = typename IO::S;
        S Uo[4], Fs[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) Uo[q] = sm.U[f.fe][f.ol][f.ok * 4 + q];
        if (f.bnd) bc_exterior(P.bc, f.dir, Uo, f.fnx, f.fny, f.Un);
        if (f.face == 0 || f.face == 2)
            riemann2(P.gas, P.riemann, f.Un, Uo, f.fnx, f.fny, Fs);
        else
            riemann2(P.gas, P.riemann, Uo, f.Un, f.fnx, f.fny, Fs);
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const S fn = sm.F[f.fe][f.ol][f.ok * 4 + q] * f.fnx + sm.G[f.fe][f.ol][f.ok * 4 + q] * f.fny;
            sm.Del[f.fe][f.face][f.idx][q] = f.sc * (Fs[q] - fn);
        }
    }

    // Volume term along one line of one element for equation q.
    //   xi-line  (row l):    U[l][k] <- xi_x d/dxi F + xi_y d/dxi G   (k = 0..N-1)
    //   eta-line (column k): V[l][k] <- eta_x d/deta F + eta_y d/deta G (l = 0..N-1)
    // The line is held in registers; D comes through the scalar cache.  Affine
    // meshes keep only d/dxi F and d/deta G (the cross terms are exact zeros).
    template <int N, class S>
    __device__ __forceinline__ void volume_line(const Res2DP& P, TileSmem<N, S>& sm, bool eta, int le, int ln,
                                                int q, long long eg)
    {
        constexpr int NN = N * N;
        const double* D  = P.tab;
        if (!eta)
        {
            S f[N], g[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                f[m] = sm.F[le][ln][m * 4 + q];
                if (P.curved) g[m] = sm.G[le][ln][m * 4 + q];
            }
            if (P.curved)
            {
                const double* mx = P.vmet + eg * 4 * NN + ln * N;
                const double* my = mx + NN;
#pragma unroll
                for (int k = 0; k < N; ++k)
                {
                    S dF = Sc<S>::c(0.0), dG = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m)
                    {
                        dF = dF + D[k * N + m] * f[m];
                        dG = dG + D[k * N + m] * g[m];
                    }
                    sm.U[le][ln][k * 4 + q] = mx[k] * dF + my[k] * dG;
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < N; ++k)
                {
                    S dF = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m) dF = dF + D[k * N + m] * f[m];
                    sm.U[le][ln][k * 4 + q] = P.xi_x * dF;
                }
            }
        }
        else
        {
            S f[N], g[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                g[m] = sm.G[le][m][ln * 4 + q];
                if (P.curved) f[m] = sm.F[le][m][ln * 4 + q];
            }
            if (P.curved)
            {
                const double* ex = P.vmet + eg * 4 * NN + 2 * NN + ln * N;
                const double* ey = ex + NN;
#pragma unroll
                for (int l = 0; l < N; ++l)
                {
                    S dF = Sc<S>::c(0.0), dG = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m)
                    {
                        dF = dF + D[l * N + m] * f[m];
                        dG = dG + D[l * N + m] * g[m];
                    }
                    sm.V[le][l][ln * 4 + q] = ex[l] * dF + ey[l] * dG;
                }
            }
            else
            {
#pragma unroll
                for (int l = 0; l < N; ++l)
                {
                    S dG = Sc<S>::c(0.0);
#pragma unroll
                    for (int m = 0; m < N; ++m) dG = dG + D[l * N + m] * g[m];
                    sm.V[le][l][ln * 4 + q] = P.eta_y * dG;
                }
            }
        }
    }

    // One tile = EPB consecutive elements of one mesh, processed by one
    // workgroup.  `epb` <= EPB elements are active (the batched sweep may run
    // smaller tiles).  With `filt` (stage outputs only) the element-local
    // spectral filter of the post-step hook is applied in LDS before the store.
    // Phases: (1) node state -> LDS, fluxes F, G -> LDS (neighbour traces of the
    // face tasks already in flight); (2) face Riemann fluxes -> Del; (3) volume
    // term by xi- and eta-lines; (4) per node: volume + VCJH lift -> output.
    template <int N, class IO>
    __device__ __forceinline__ void res2d_tile(const Res2DP& P, const IO& io, long long blk, int epb, bool filt,
                                               TileSmem<N, typename IO::S>& sm)
    {
        using S           = typename IO::S;
        constexpr int BS  = Tile2D<N, S>::BS;
        constexpr int NN  = N * N;
        const int     tid = threadIdx.x;
        if (filt)
            for (int q = tid; q < NN; q += BS) sm.Fm[q] = P.tab[NN + 2 * N + q];
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

        const TileOrigin org = tile_origin(P, blk, epb);
        FaceTask<N, S>   f0;
        face_task<N, IO>(P, io, blk, epb, org, tid, f0);

        S own[4];
        if (act)
        {
            io.ld4((e * NN + node) * 4, own);
            S F[4], G[4];
            flux2(P.gas, own, F, G);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                sm.U[el][l][k * 4 + q] = own[q];
                sm.F[el][l][k * 4 + q] = F[q];
                sm.G[el][l][k * 4 + q] = G[q];
            }
        }
        __syncthreads();

        // ---- face tasks: 4 faces x N nodes per element ----
        if (f0.live) face_flux<N, IO>(P, f0, sm);
        for (int t = tid + BS; t < epb * 4 * N; t += BS)
        {
            FaceTask<N, S> f;
            face_task<N, IO>(P, io, blk, epb, org, t, f);
            if (f.live) face_flux<N, IO>(P, f, sm);
        }
        __syncthreads();

        // ---- volume term by lines (q fastest).  Curved: xi-line tasks, then
        // eta-line tasks (4 contractions per node); affine: one task does
        // xi-line ln of F and eta-line ln of G (2 contractions per node) ----
        {
            const int nlt = epb * N * 4;
            const int nt  = P.curved ? 2 * nlt : nlt;
            for (int t = tid; t < nt; t += BS)
            {
                const bool      eta = t >= nlt;
                const int       w   = eta ? t - nlt : t;
                const int       q   = w & 3;
                const int       le  = (w >> 2) / N;
                const int       ln  = (w >> 2) - le * N;
                const long long eg  = blk * epb + le;
                if (eg >= P.E) continue;
                if (P.curved)
                    volume_line<N, S>(P, sm, eta, le, ln, q, eg);
                else
                {
                    volume_line<N, S>(P, sm, false, le, ln, q, eg);
                    volume_line<N, S>(P, sm, true, le, ln, q, eg);
                }
            }
        }
        __syncthreads();

        S r[4];
        if (act)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                r[q] = sm.U[el][l][k * 4 + q] + sm.V[el][l][k * 4 + q];
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
                // stage value -> LDS (each node overwrites only its own slot of
                // U, which only this thread reads in this phase), then
                // out = 1/2 (F_x v + F_y v)
                if (act)
                {
                    double o[4];
                    io.val4((e * NN + node) * 4, own, r, o);
#pragma unroll
                    for (int q = 0; q < 4; ++q) sm.U[el][l][k * 4 + q] = o[q];
                }
                __syncthreads();
                if (act)
                {
                    double o[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                    {
                        double sx = 0.0, sy = 0.0;
                        for (int c = 0; c < N; ++c)
                        {
                            sx += sm.Fm[k * N + c] * sm.U[el][l][c * 4 + q];
                            sy += sm.Fm[l * N + c] * sm.U[el][c][k * 4 + q];
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
#pragma unroll
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
#pragma unroll
            for (int q = 0; q < 3; ++q) prev[q] = io.ld(static_cast<long long>(ep) * 3 * N + q * N + (N - 1));
        if (e == P.ne - 1 && !P.periodic)
            bc_exterior1(P.bc, 1, last, next);
        else
#pragma unroll
            for (int q = 0; q < 3; ++q) next[q] = io.ld(static_cast<long long>(en) * 3 * N + q * N);
        S res[3][N];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            for (int m = 0; m < N; ++m) res[q][m] = Sc<S>::c(0.0);
        for (int jn = 0; jn < N; ++jn)
        {
            S nd[3] = {u[0][jn], u[1][jn], u[2][jn]}, F[3];
            flux1(P.gas, nd, F);
            for (int in = 0; in < N; ++in)
            {
                const double Dij = P.inv_jac * D[in * N + jn];
#pragma unroll
                for (int q = 0; q < 3; ++q) res[q][in] = res[q][in] + Dij * F[q];
            }
        }
        S FdL[3], FdR[3], FsL[3], FsR[3];
        flux1(P.gas, first, FdL);
        flux1(P.gas, last, FdR);
        riemann1(P.gas, P.riemann, prev, first, FsL);
        riemann1(P.gas, P.riemann, last, next, FsR);
#pragma unroll
        for (int q = 0; q < 3; ++q)
        {
            const S dl = FsL[q] - FdL[q], dr = FsR[q] - FdR[q];
            for (int in = 0; in < N; ++in)
                res[q][in] = res[q][in] + P.inv_jac * (dl * dL[in] + dr * dR[in]);
        }
#pragma unroll
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
#pragma unroll
            for (int q = 0; q < 4; ++q) sU[el][node][q] = o[q];
        }
        __syncthreads();
        if (!act) return;
        double o[4];
#pragma unroll
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
#pragma unroll
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < N; ++r)
            {
                double s = 0.0;
                for (int c = 0; c < N; ++c) s += Fm[transpose ? c * N + r : r * N + c] * x[q * N + c];
                out[static_cast<long long>(e) * 3 * N + q * N + r] = s;
            }
    }
} // namespace frsdg

// frsdg_adjoint.hpp — adjoint (vector-Jacobian) kernels: out = sigma*w + (dR/du)^T w.
//
// R of element e is   V_e(u_e) + sum over its 4 face copies of  lift_f (Delta_f)
// with Delta_f = s_f (F*(U_L, U_R; n_f) - F(U_own).n_f).  Transposing:
//   k_adj_lift: per element and face copy, the lifted cotangent
//     Wbar_{e,f}(q, idx) = sum_k W_e(q, line(idx, k)) g'(k)      (g' = dL or dR);
//   k_adj2d_tile: per tile of elements (the residual's tiles),
//     volume: Fbar(q,k,l) = sum_m D[m][k] (xi_x W)(q,m,l) + sum_m D[m][l] (eta_x W)(q,k,m),
//             Gbar likewise with xi_y, eta_y;  ubar += J_F^T Fbar + J_G^T Gbar per node;
//     faces : each face once: ubar_A, ubar_B += J_{F*}^T s (Wbar_A + Wbar_B) for both
//             traces, minus s J_{Fn,own}^T Wbar_own per side; boundary faces chain
//             through the BC map.
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

    // lifted cotangents of the face copies: Wbar[e][f][idx][q].  One thread
    // per element line: row idx gives the x faces 0 (with dL) and 1 (dR),
    // column idx the y faces 2 and 3, from one read of the line.
    template <int N>
    __global__ __launch_bounds__(256) void k_adj_lift(long long E, const double* __restrict__ tab,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ wb)
    {
        constexpr int NN = N * N;
        const long long t = blockIdx.x * 256LL + threadIdx.x;
        if (t >= E * 2 * N) return;
        const long long e   = t / (2 * N);
        const int       pt  = static_cast<int>(t - e * 2 * N);
        const int       yf  = pt / N, idx = pt - yf * N; // yf: column (y faces) or row (x faces)
        const double*   gL  = tab + NN;
        const double*   gR  = tab + NN + N;
        const double*   we  = w + e * NN * 4;
        double          a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < N; ++k)
        {
            const int    node = yf ? (k * N + idx) : (idx * N + k);
            const double2 v0  = *reinterpret_cast<const double2*>(we + node * 4);
            const double2 v1  = *reinterpret_cast<const double2*>(we + node * 4 + 2);
            const double  x[4] = {v0.x, v0.y, v1.x, v1.y};
            const double  l = gL[k], r = gR[k];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                a0[q] += x[q] * l;
                a1[q] += x[q] * r;
            }
        }
        double* o = wb + ((e * 4 + 2 * yf) * N + idx) * 4; // face 2 yf, then face 2 yf + 1 (N entries on)
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            o[q]         = a0[q];
            o[N * 4 + q] = a1[q];
        }
    }

    // ------------------------------------------------------------------
    // Tile adjoint (2D): the residual's tiles (256 threads, TX x TY elements,
    // one thread per node), each face of a tile evaluated once.
    //   load     : u, w -> LDS;
    //   phase 1  : xi-line tasks  Fbar(l, k) = sum_m D[m][k] xi_x W(l, m),
    //              Gbar likewise with xi_y (affine: eta-line tasks Gbar from
    //              eta_y at the same time), and face tasks: one reverse Riemann
    //              sweep per face node with cs = s (Wbar_A + Wbar_B) gives the
    //              cotangents of both traces (rim faces keep their own side,
    //              the neighbouring tile evaluates the other), plus the own
    //              normal-flux term -s Wbar through flux2_vjp; results -> Cs;
    //   phase 2  : (curved) eta-line tasks add sum_m D[m][l] (eta_x, eta_y) W(m, k);
    //   per node : ubar = (dF/dU)^T Fbar + (dG/dU)^T Gbar + the face slots,
    //              out = AdjOut(w, ubar).
    // Wbar (the lifted cotangents of every element's face copies) comes from
    // k_adj_lift; periodic wraps whose two copies carry different metrics
    // evaluate the outer copy's Riemann flux a second time with its metrics.
    // ------------------------------------------------------------------
    template <int N>
    struct AdjSmem
    {
        static constexpr int EPB = Tile2D<N, double>::EPB;
        static constexpr int RB  = N * 4 * 8;
        static constexpr int RP  = N * 4 + ((RB % 64) == 0 ? 4 : 0);
        __device__ static __forceinline__ int pos(int l, int k, int q) { return 4 * k + (q ^ (l & 3)); }
        double    U[EPB][N][RP];    // state
        double    W[EPB][N][RP];    // cotangent
        double    FX[EPB][N][RP];   // Fbar
        double    GX[EPB][N][RP];   // Gbar
        double    Cs[EPB][4][N][4]; // face cotangents of the own trace (faces 0 left, 1 right, 2 bottom, 3 top)
        long long eid[EPB];
    };

    struct AdjFace
    {
        bool   live, xdir, wrap;
        int    idx, elA, elB, bdir;
        double fnx, fny, sc;    // metrics of the face index this tile evaluates
        double nx1, ny1, sc1;   // wrap with differing ends: the outer element's own copy
        double Uext[4];         // outer trace
        double wA[4], wB[4];    // lifted cotangents of side A's / side B's copy (0 beyond a physical boundary)
    };

    __device__ __forceinline__ void face_metrics(const Res2DP& P, bool xdir, int j, int i, int idx, int N,
                                                 double& nx, double& ny, double& sc)
    {
        if (!P.curved)
        {
            nx = xdir ? 1.0 : 0.0;
            ny = xdir ? 0.0 : 1.0;
            sc = xdir ? P.sc_x : P.sc_y;
            return;
        }
        // x faces: face index i in [0, nx] of row j; y faces: face index j in [0, ny] of column i
        const double* fm = xdir ? P.fxm + ((static_cast<long long>(j) * (P.nx + 1) + i) * N + idx) * 4
                                : P.fym + ((static_cast<long long>(j) * P.nx + i) * N + idx) * 4;
        nx = fm[0];
        ny = fm[1];
        sc = fm[2];
    }

    // geometry + HBM gathers of face task t of a tile (same enumeration as face_task)
    template <int N>
    __device__ __forceinline__ void adj_face_setup(const Res2DP& P, const double* __restrict__ u,
                                                   const double* __restrict__ wb, const TileGeom& g, int t,
                                                   AdjFace& f)
    {
        constexpr int NN = N * N;
        const int     XT = (g.tx + 1) * g.ty * N;
        const int     YT = g.tx * (g.ty + 1) * N;
        f.live           = false;
        if (t >= XT + YT) return;
        f.xdir = t < XT;
        f.bdir = -1;
        f.wrap = false;
        long long eA = -1, eB = -1; // global elements of the two sides (-1: physical boundary)
        int       fo = 0, io = 0, jo = 0; // the outer element's own face index (wrap faces)
        bool      outerA = false, wrapped = false;
        long long eo   = -1;
        int       node = 0;
        if (f.xdir)
        {
            const int b  = t / ((g.tx + 1) * N);
            const int r  = t - b * (g.tx + 1) * N;
            const int a  = r / N;
            f.idx        = r - a * N;
            const int j  = g.j0 + b;
            const int fi = g.i0 + a;
            if (j >= P.ny) return;
            const bool inA = a >= 1 && fi - 1 < P.nx;
            const bool inB = a < g.tx && fi < P.nx;
            if (!inA && !inB) return;
            f.live = true;
            f.elA  = inA ? b * g.tx + a - 1 : -1;
            f.elB  = inB ? b * g.tx + a : -1;
            face_metrics(P, true, j, fi, f.idx, N, f.fnx, f.fny, f.sc);
            if (inA) eA = static_cast<long long>(j) * P.nx + fi - 1;
            if (inB) eB = static_cast<long long>(j) * P.nx + fi;
            if (!inA)
            {
                outerA = true;
                int ni = fi - 1;
                if (ni < 0)
                {
                    if (P.per_x)
                    {
                        ni      = P.nx - 1;
                        wrapped = true;
                        fo      = P.nx;
                    }
                    else
                        f.bdir = 0;
                }
                if (f.bdir < 0)
                {
                    eo   = static_cast<long long>(j) * P.nx + ni;
                    eA   = eo;
                    node = f.idx * N + (N - 1);
                }
                jo = j;
            }
            else if (!inB)
            {
                int ni = fi;
                if (ni >= P.nx)
                {
                    if (P.per_x)
                    {
                        ni      = 0;
                        wrapped = true;
                        fo      = 0;
                    }
                    else
                        f.bdir = 1;
                }
                if (f.bdir < 0)
                {
                    eo   = static_cast<long long>(j) * P.nx + ni;
                    eB   = eo;
                    node = f.idx * N;
                }
                jo = j;
            }
            io = fo;
        }
        else
        {
            const int uu = t - XT;
            const int bf = uu / (g.tx * N);
            const int r  = uu - bf * g.tx * N;
            const int a  = r / N;
            f.idx        = r - a * N;
            const int i  = g.i0 + a;
            const int fj = g.j0 + bf;
            if (i >= P.nx) return;
            const bool inA = bf >= 1 && fj - 1 < P.ny;
            const bool inB = bf < g.ty && fj < P.ny;
            if (!inA && !inB) return;
            f.live = true;
            f.elA  = inA ? (bf - 1) * g.tx + a : -1;
            f.elB  = inB ? bf * g.tx + a : -1;
            face_metrics(P, false, fj, i, f.idx, N, f.fnx, f.fny, f.sc);
            if (inA) eA = static_cast<long long>(fj - 1) * P.nx + i;
            if (inB) eB = static_cast<long long>(fj) * P.nx + i;
            if (!inA)
            {
                outerA = true;
                int nj = fj - 1;
                if (nj < 0)
                {
                    if (P.per_y)
                    {
                        nj      = P.ny - 1;
                        wrapped = true;
                        fo      = P.ny;
                    }
                    else
                        f.bdir = 3;
                }
                if (f.bdir < 0)
                {
                    eo   = static_cast<long long>(nj) * P.nx + i;
                    eA   = eo;
                    node = (N - 1) * N + f.idx;
                }
            }
            else if (!inB)
            {
                int nj = fj;
                if (nj >= P.ny)
                {
                    if (P.per_y)
                    {
                        nj      = 0;
                        wrapped = true;
                        fo      = 0;
                    }
                    else
                        f.bdir = 2;
                }
                if (f.bdir < 0)
                {
                    eo   = static_cast<long long>(nj) * P.nx + i;
                    eB   = eo;
                    node = f.idx;
                }
            }
            io = i;
            jo = fo;
        }
        if (eo >= 0) ld4d(u + (eo * NN + node) * 4, f.Uext);
        const int fa = f.xdir ? 1 : 3, fb = f.xdir ? 0 : 2;
        if (eA >= 0)
            ld4d(wb + ((eA * 4 + fa) * N + f.idx) * 4, f.wA);
        else
#pragma unroll
            for (int q = 0; q < 4; ++q) f.wA[q] = 0.0;
        if (eB >= 0)
            ld4d(wb + ((eB * 4 + fb) * N + f.idx) * 4, f.wB);
        else
#pragma unroll
            for (int q = 0; q < 4; ++q) f.wB[q] = 0.0;
        if (wrapped && P.curved)
        {
            face_metrics(P, f.xdir, jo, io, f.idx, N, f.nx1, f.ny1, f.sc1);
            f.wrap = !(f.nx1 == f.fnx && f.ny1 == f.fny && f.sc1 == f.sc);
        }
        (void)outerA;
    }

    template <int N>
    __device__ __forceinline__ void adj_face_eval(const Res2DP& P, const AdjFace& f, AdjSmem<N>& sm)
    {
        const int kA = f.xdir ? N - 1 : f.idx, lA = f.xdir ? f.idx : N - 1;
        const int kB = f.xdir ? 0 : f.idx, lB = f.xdir ? f.idx : 0;
        double    UA[4], UB[4];
        if (f.elA >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) UA[q] = sm.U[f.elA][lA][sm.pos(lA, kA, q)];
        }
        if (f.elB >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) UB[q] = sm.U[f.elB][lB][sm.pos(lB, kB, q)];
        }
        if (f.elA < 0)
        {
            if (f.bdir >= 0)
                bc_exterior(P.bc, f.bdir, UB, f.fnx, f.fny, UA);
            else
#pragma unroll
                for (int q = 0; q < 4; ++q) UA[q] = f.Uext[q];
        }
        if (f.elB < 0)
        {
            if (f.bdir >= 0)
                bc_exterior(P.bc, f.bdir, UA, f.fnx, f.fny, UB);
            else
#pragma unroll
                for (int q = 0; q < 4; ++q) UB[q] = f.Uext[q];
        }
        // The cotangents of the Riemann flux (both copies of the face; the outer
        // copy of a wrap with differing ends separately) are formed first, and
        // the own-normal-flux parts -s Wbar . n of the tile sides go to the
        // face slots of Cs (each slot belongs to this face task alone) before
        // the Riemann sweep, whose part is added afterwards: neither the lifted
        // cotangents wA, wB nor those parts are held in registers across it.
        double cs[4], c2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            if (!f.wrap)
                cs[q] = f.sc * f.wA[q] + f.sc * f.wB[q];
            else
            {
                cs[q] = f.sc * (f.elA >= 0 ? f.wA[q] : f.wB[q]);
                c2[q] = f.sc1 * (f.elA >= 0 ? f.wB[q] : f.wA[q]);
            }
        }
        const int fa = f.xdir ? 1 : 3, fb = f.xdir ? 0 : 2;
        if (f.elA >= 0)
        {
            double Fb[4], Gb[4], b[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double cn = -f.sc * f.wA[q];
                Fb[q]           = cn * f.fnx;
                Gb[q]           = cn * f.fny;
            }
            flux2_vjp(P.gas, UA, Fb, Gb, b);
#pragma unroll
            for (int q = 0; q < 4; ++q) sm.Cs[f.elA][fa][f.idx][q] = b[q];
        }
        if (f.elB >= 0)
        {
            double Fb[4], Gb[4], b[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double cn = -f.sc * f.wB[q];
                Fb[q]           = cn * f.fnx;
                Gb[q]           = cn * f.fny;
            }
            flux2_vjp(P.gas, UB, Fb, Gb, b);
#pragma unroll
            for (int q = 0; q < 4; ++q) sm.Cs[f.elB][fb][f.idx][q] = b[q];
        }
        double bA[4] = {0, 0, 0, 0}, bB[4] = {0, 0, 0, 0};
        riemann2_vjp(P.gas, P.riemann, UA, UB, f.fnx, f.fny, cs, bA, bB);
        if (f.wrap) riemann2_vjp(P.gas, P.riemann, UA, UB, f.nx1, f.ny1, c2, bA, bB);
        // physical boundary: the exterior state is a function of the own trace
        if (f.bdir >= 0)
        {
            if (f.elA < 0)
                bc_vjp(P.bc, f.bdir, UB, f.fnx, f.fny, bA, bB);
            else
                bc_vjp(P.bc, f.bdir, UA, f.fnx, f.fny, bB, bA);
        }
        if (f.elA >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) sm.Cs[f.elA][fa][f.idx][q] += bA[q];
        }
        if (f.elB >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) sm.Cs[f.elB][fb][f.idx][q] += bB[q];
        }
    }

    // transposed D contraction along one line of element le, equation q:
    //   xi-line  (row ln):    FX[ln][k] (+)= sum_m D[m][k] ma[m] W[ln][m], GX likewise with mb
    //   eta-line (column ln): FX[l][ln] (+)= sum_m D[m][l] ma[m] W[m][ln], GX likewise with mb
    // (ma, mb) = (xi_x, xi_y) on xi-lines, (eta_x, eta_y) on eta-lines; affine:
    // xi-lines write FX from xi_x only, eta-lines GX from eta_y only.  D^T is
    // centro-antisymmetric like D: the even-odd form (eo_line) with its own
    // table.  Curved tiles add both lines' parts into FX, GX (zeroed per node
    // first) by LDS atomics: two addends per entry, fl(xi + eta) whichever
    // lands first.
    template <int N, bool CURVED>
    __device__ __forceinline__ void adj_line(const Res2DP& P, AdjSmem<N>& sm, bool eta, int le, int ln, int q)
    {
        const double* T = eo_tab<N>(P.tab, 2);
        double        a[N], b[N], f[N], g[N];
        if constexpr (CURVED)
        {
            double ma[N], mb[N];
            line_metrics<N>(P.vmet, sm.eid[le], eta, ln, ma, mb);
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                const double w = eta ? sm.W[le][m][sm.pos(m, ln, q)] : sm.W[le][ln][sm.pos(ln, m, q)];
                a[m]           = ma[m] * w;
                b[m]           = mb[m] * w;
            }
            eo_line<N, double, true>(T, a, f);
            eo_line<N, double, true>(T, b, g);
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                const int l = eta ? r : ln, k = eta ? ln : r;
                atomicAdd(&sm.FX[le][l][sm.pos(l, k, q)], f[r]);
                atomicAdd(&sm.GX[le][l][sm.pos(l, k, q)], g[r]);
            }
        }
        else
        {
            const double mt = eta ? P.eta_y : P.xi_x;
#pragma unroll
            for (int m = 0; m < N; ++m)
                a[m] = mt * (eta ? sm.W[le][m][sm.pos(m, ln, q)] : sm.W[le][ln][sm.pos(ln, m, q)]);
            eo_line<N, double, true>(T, a, f);
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                const int l = eta ? r : ln, k = eta ? ln : r;
                if (eta)
                    sm.GX[le][l][sm.pos(l, k, q)] = f[r];
                else
                    sm.FX[le][l][sm.pos(l, k, q)] = f[r];
            }
        }
    }

    // Register budget of the adjoint tiles: 168 VGPRs (three waves per SIMD,
    // three tiles per CU, which the LDS of n >= 3 allows) for 8-18 spilled
    // registers in the reverse Riemann sweep, against 180-190 VGPRs and two
    // waves unconstrained: measured 17 % faster on warped 1024^2 p=7 (4.44 ->
    // 3.74 ms) and 11 % on the 128^2 p=4 vortex.  n = 2 tiles (66 KB of LDS,
    // two per CU) stay unconstrained.
    template <int N>
    constexpr int adj_waves()
    {
        return N >= 3 ? 3 : 1;
    }

    template <int N, bool CURVED>
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(adj_waves<N>())))
    void k_adj2d_tile(const Res2DP P, const double* __restrict__ u, const double* __restrict__ w,
                      const double* __restrict__ wb, const AdjOut ao, double* __restrict__ out)
    {
        using T           = Tile2D<N, double>;
        constexpr int BS  = 256;
        constexpr int NN  = N * N;
        __shared__ AdjSmem<N> sm;
        const int      tid  = threadIdx.x;
        const TileGeom g    = tile_geom(P, xcd_tile(blockIdx.x, gridDim.x), T::TX, T::TY);
        const int      tx = T::TX, ty = T::TY;
        const int      el   = tid / NN;
        const int      node = tid - el * NN;
        const int      k    = node % N;
        const int      l    = node / N;
        const int      eb   = el / tx;
        const int      ea   = el - eb * tx;
        const bool     act  = (el < tx * ty) && (g.i0 + ea < P.nx) && (g.j0 + eb < P.ny);
        const long long e   = static_cast<long long>(g.j0 + eb) * P.nx + (g.i0 + ea);
        if (node == 0 && el < tx * ty) sm.eid[el] = act ? e : -1;

        // tasks: xi- and eta-line tasks, then face tasks
        const int nlt = tx * ty * N * 4;
        const int nl  = 2 * nlt;
        const int nft = ((tx + 1) * ty + tx * (ty + 1)) * N;

        if (act)
        {
            double U[4], Wv[4];
            ld4d(u + (e * NN + node) * 4, U);
            ld4d(w + (e * NN + node) * 4, Wv);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                sm.U[el][l][sm.pos(l, k, q)] = U[q];
                sm.W[el][l][sm.pos(l, k, q)] = Wv[q];
                if constexpr (CURVED)
                {
                    sm.FX[el][l][sm.pos(l, k, q)] = 0.0;
                    sm.GX[el][l][sm.pos(l, k, q)] = 0.0;
                }
            }
        }
        __syncthreads();

        // one phase: face tasks and line tasks (xi and eta; curved: added into
        // FX, GX) read U and W and write disjoint arrays.  Where the face tasks
        // fit in waves 0-1 and the line tasks twice in waves 2-3 (curved n = 8)
        // the two run on separate waves; otherwise line tasks first, then faces
        // (two loops: one loop over both kinds holds a line task's registers
        // across the face sweep, 250 VGPRs against 190).
        if (CURVED && nft <= BS / 2 && nl <= BS)
        {
            if (tid < BS / 2)
            {
                if (tid < nft)
                {
                    AdjFace f;
                    adj_face_setup<N>(P, u, wb, g, tid, f);
                    if (f.live) adj_face_eval<N>(P, f, sm);
                }
            }
            else
            {
                for (int t = tid - BS / 2; t < nl; t += BS / 2)
                {
                    const bool eta = t >= nlt;
                    const int  wq  = eta ? t - nlt : t;
                    const int  le  = (wq >> 2) / N;
                    if (sm.eid[le] >= 0) adj_line<N, CURVED>(P, sm, eta, le, (wq >> 2) - le * N, wq & 3);
                }
            }
        }
        else
        {
            for (int t = tid; t < nl; t += BS)
            {
                const bool eta = t >= nlt;
                const int  wq  = eta ? t - nlt : t;
                const int  le  = (wq >> 2) / N;
                if (sm.eid[le] < 0) continue;
                adj_line<N, CURVED>(P, sm, eta, le, (wq >> 2) - le * N, wq & 3);
            }
            for (int t = tid; t < nft; t += BS)
            {
                AdjFace f;
                adj_face_setup<N>(P, u, wb, g, t, f);
                if (f.live) adj_face_eval<N>(P, f, sm);
            }
        }
        __syncthreads();
        if (!act) return;

        double U[4], Wv[4], Fb[4], Gb[4], ub[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            U[q]  = sm.U[el][l][sm.pos(l, k, q)];
            Wv[q] = sm.W[el][l][sm.pos(l, k, q)];
            Fb[q] = sm.FX[el][l][sm.pos(l, k, q)];
            Gb[q] = sm.GX[el][l][sm.pos(l, k, q)];
        }
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
                for (int q = 0; q < 4; ++q) acc += sr[q].d * Wv[q];
                ub[i] += acc;
            }
        }
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            double s = ub[q];
            if (k == 0) s += sm.Cs[el][0][l][q];
            if (k == N - 1) s += sm.Cs[el][1][l][q];
            if (l == 0) s += sm.Cs[el][2][k][q];
            if (l == N - 1) s += sm.Cs[el][3][k][q];
            o[q] = ao.at((e * NN + node) * 4 + q, Wv[q], s);
        }
        st4d(out + (e * NN + node) * 4, o);
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

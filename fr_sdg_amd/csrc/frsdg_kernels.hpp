// frsdg_kernels.hpp — HIP kernels of the FR residual path for gfx950.
//
// Residual kernel design (2D, one launch per residual or SSP-RK stage):
//   * a workgroup owns a TX x TY block (tile) of the element grid, one thread
//     per solution node (EPB = floor(256 / n^2) elements; n = 8: 2 x 2); the
//     element blocks are contiguous in HBM, so the 32-byte node loads / stores
//     of a workgroup are coalesced;
//   * each thread evaluates the Euler fluxes F, G of its node once and stages
//     U, F, G in LDS;
//   * face tasks: every face of the tile is evaluated once.  Interior faces
//     take both traces from LDS; faces on the tile rim gather the outer trace
//     from HBM (issued at kernel start) or form it from the boundary
//     condition.  The scaled flux jump is written for the tile elements on
//     both sides.  Neighbouring tiles evaluate a shared rim face with
//     identical arguments: bitwise the same flux, conservative and
//     deterministic, no atomics;
//   * the derivative-matrix contraction runs as line tasks (one element row
//     or column, one equation per thread): the line is read from LDS once into
//     registers, D comes through the scalar cache, and the metric-weighted
//     derivatives of all n nodes of the line go back to LDS (curved double
//     tiles: added by LDS atomics into one zeroed array, so that the line
//     tasks need not wait for the face tasks; where the face tasks fit in two
//     wavefronts, as at n = 8, those two evaluate the faces while the other
//     two run the line tasks);
//   * per node: volume term + the four VCJH lifts, written once; output
//     policies fuse rhs = -R, F = udot + R and the SSP-RK stage update (the
//     last stage of a step optionally applies the post-step spectral filter
//     in LDS before the store) into the same store;
//   * affine meshes use the four constant metric terms (only d/dxi of F and
//     d/deta of G survive, as in the reference where the cross terms are exact
//     zeros); curved meshes read per-node metrics and per-face (nx, ny, scale)
//     precomputed once in HBM;
//   * block index -> tile is remapped so that each XCD walks a contiguous band
//     of tiles, a vertical strip of the tile grid walked row by row (neighbour
//     traces hit its own L2);
//   * the output's extra operand (the stage's u^n, the ifunction's udot) is
//     loaded before the last barrier, in flight across the lift phase.
// The same tile code is instantiated on double (residual, stages) and Cplx
// (complex step, frsdg_scalar.hpp).  The forward tangent J*v (k_tan2d) runs
// the same tiles with forward-mode pairs only in the node and face fluxes.
#pragma once
#include <type_traits>
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
        const double* __restrict__ rinv;   // [E][n*n] 1/radius of radial meshes, else nullptr
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

    // Health of a stored 2D node state (rho, rho u, rho v, rho E): bit 1 a
    // non-finite value, bit 2 rho <= 0, bit 4 internal energy <= 0
    // (2 rho (rho E) <= |rho v|^2).  Flags are OR-ed into *flag, only on failure.
    __device__ __forceinline__ void health_check(const double o[4], int* flag)
    {
        int f = 0;
        if (!__builtin_isfinite(o[0] + o[1] + o[2] + o[3])) f |= 1;
        if (!(o[0] > 0.0)) f |= 2;
        if (!(2.0 * o[0] * o[3] > o[1] * o[1] + o[2] * o[2])) f |= 4;
        if (f) atomicOr(flag, f);
    }

    // out = add + scale * R   (residual: scale 1; rhs: scale -1; ifunction: add = udot);
    // with un (and add null): out = shift (u - un) + R, the implicit function of
    // a backward-Euler stage with its time derivative formed in the store
    struct IORes
    {
        using S = double;
        static constexpr bool can_filter = false;
        const double* __restrict__ u;
        double* __restrict__ out;
        const double* __restrict__ add;
        double scale;
        const double* __restrict__ un = nullptr;
        double shift                  = 0.0;
        __device__ double ld(long long i) const { return u[i]; }
        __device__ void ld4(long long i, double o[4]) const { ld4d(u + i, o); }
        __device__ void st(long long i, double, double r) const
        {
            out[i] = add ? add[i] + r : scale * r;
        }
        // the added operand of node i, loaded ahead of the last phase (st4p)
        static constexpr bool pre = true;
        __device__ void pre4(long long i, double z[4]) const
        {
            if (add)
                ld4d(add + i, z);
            else if (un)
                ld4d(un + i, z);
        }
        __device__ void st4p(long long i, const double* own, const double r[4], const double z[4]) const
        {
            double o[4];
            if (add)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = z[q] + r[q];
            }
            else if (un)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = __builtin_fma(shift, own[q] - z[q], r[q]);
            }
            else
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = scale * r[q];
            st4d(out + i, o);
        }
        __device__ void st4(long long i, const double* own, const double r[4]) const
        {
            double z[4];
            pre4(i, z);
            st4p(i, own, r, z);
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
        int*    health = nullptr; // 2D: flags of the stored node states (health_flags), may be null
        __device__ double ld(long long i) const { return u[i]; }
        __device__ void ld4(long long i, double o[4]) const { ld4d(u + i, o); }
        __device__ void st(long long i, double own, double r) const
        {
            const double s = __builtin_fma(-cdt, r, own);
            out[i]         = (a0 != 0.0) ? __builtin_fma(b, s, a0 * u0[i]) : b * s;
        }
        // u0 of node i, loaded ahead of the last phase (val4p / st4p)
        static constexpr bool pre = true;
        __device__ void pre4(long long i, double z[4]) const
        {
            if (a0 != 0.0) ld4d(u0 + i, z);
        }
        // the contractions are written out, so that every kernel that stores
        // a stage rounds it the same way
        __device__ void val4p(const double own[4], const double r[4], const double z[4], double o[4]) const
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = __builtin_fma(-cdt, r[q], own[q]);
            if (a0 != 0.0)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = __builtin_fma(b, o[q], a0 * z[q]);
            }
            else
#pragma unroll
                for (int q = 0; q < 4; ++q) o[q] = b * o[q];
        }
        __device__ void val4(long long i, const double own[4], const double r[4], double o[4]) const
        {
            double z[4];
            pre4(i, z);
            val4p(own, r, z, o);
        }
        __device__ void put4(long long i, const double o[4]) const
        {
            if (health) health_check(o, health);
            st4d(out + i, o);
        }
        __device__ void st4p(long long i, const double own[4], const double r[4], const double z[4]) const
        {
            double o[4];
            val4p(own, r, z, o);
            put4(i, o);
        }
        __device__ void st4(long long i, const double own[4], const double r[4]) const
        {
            double o[4];
            val4(i, own, r, o);
            put4(i, o);
        }
    };
    // forward tangent: out = sigma*v + dR
    struct IOTan
    {
        using S = Dual;
        static constexpr bool can_filter = false;
        static constexpr bool pre        = false;
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
        static constexpr bool pre        = false;
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
    struct BlockSize<Cplx>
    {
        static constexpr int value = 128;
    };
    // largest divisor of epb that is <= sqrt(epb): rows of a 2D element tile
    __host__ __device__ constexpr int tile_rows(int epb)
    {
        int best = 1;
        for (int d = 1; d * d <= epb; ++d)
            if (epb % d == 0) best = d;
        return best;
    }
    // Elements of one workgroup: EPB = floor(BS / n^2) elements as a TX x TY
    // block of the element grid (TX >= TY).
    template <int N, class S>
    struct Tile2D
    {
        static constexpr int BS  = BlockSize<S>::value;
        static constexpr int NN  = N * N;
        static constexpr int EPB = (BS / NN) > 0 ? (BS / NN) : 1;
        static constexpr int TY  = tile_rows(EPB);
        static constexpr int TX  = EPB / TY;
    };
    // element tile shape of order n in the double kernels: the default TX x TY
    // tile with its rows (then columns) halved log2(div) times
    __host__ __device__ constexpr int tile_halve(int tx, int ty, int div, bool want_tx)
    {
        for (int d = div; d > 1; d /= 2)
        {
            if (ty > 1)
                ty = (ty + 1) / 2;
            else if (tx > 1)
                tx = (tx + 1) / 2;
        }
        return want_tx ? tx : ty;
    }
    // element tile of order n, shape choice s: 0 the default tile, 1 the
    // default tile halved (tile_halve with div 2), 2 the tile other than the
    // default with the fewest rim faces per element, (tx + ty) / (tx ty), over
    // tx >= ty, tx ty <= EPB (n = 7: 2 x 2 instead of 5 x 1; n = 5: 3 x 3
    // instead of 5 x 2)
    __host__ __device__ constexpr int batch_tile_dim(int n, int s, bool want_tx)
    {
        const int epb = (256 / (n * n)) > 0 ? 256 / (n * n) : 1;
        const int ty0 = tile_rows(epb), tx0 = epb / ty0;
        if (s == 1) return tile_halve(tx0, ty0, 2, want_tx);
        if (s == 2)
        {
            int bx = tx0, by = ty0, num = -1, den = 1;
            for (int ty = 1; ty * ty <= epb; ++ty)
                for (int tx = ty; tx * ty <= epb; ++tx)
                {
                    if (tx == tx0 && ty == ty0) continue;
                    // (tx + ty) / (tx ty) < num / den, ties to the larger tile
                    const long a = static_cast<long>(tx + ty) * den, b = static_cast<long>(num) * tx * ty;
                    if (num < 0 || a < b || (a == b && tx * ty > bx * by))
                    {
                        bx  = tx;
                        by  = ty;
                        num = tx + ty;
                        den = tx * ty;
                    }
                }
            return want_tx ? bx : by;
        }
        return want_tx ? tx0 : ty0;
    }
    inline void tile_shape(int n, int div, int& tx, int& ty)
    {
        const int epb = (256 / (n * n)) > 0 ? 256 / (n * n) : 1;
        const int ty0 = tile_rows(epb), tx0 = epb / ty0;
        tx            = tile_halve(tx0, ty0, div, true);
        ty            = tile_halve(tx0, ty0, div, false);
    }
    inline long long tile_count(int nx, int ny, int tx, int ty)
    {
        return static_cast<long long>((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
    }

    // ------------------------------------------------------------------
    // 2D residual: physics_impl_2D.hpp:267-357 (volume) + :689-899 (faces)
    // ------------------------------------------------------------------
    // LDS of one residual tile.  Node (k, l) of tile element el, equation q,
    // lives at X[el][l][k * 4 + q]; the arrays that line tasks read or write
    // by rows are padded so that the row stride is an odd multiple of 32
    // bytes (a wavefront of line tasks then spreads over all LDS banks).  V is
    // only written by columns and read per node: no padding (this keeps the
    // curved n = 8 tile at 39 KB, four workgroups per CU).
    //
    // Within a row the four equations of node k are stored in the order
    // q ^ (l & 3) (row l): a half-wave of per-node accesses (one q, rows
    // l..l+3) then spreads over twice to four times as many banks (n = 8:
    // 32 distinct 8-byte bank pairs instead of 8, measured 3.5 against 8.9
    // cycles per wave read), while row-line tasks (fixed k) still read
    // base + constant offsets and column-line tasks one of four bases.
    template <int N, class S, bool CURVED>
    struct TileSmem
    {
        static constexpr int NN = N * N, EPB = Tile2D<N, S>::EPB;
        static constexpr int RB = N * 4 * static_cast<int>(sizeof(S)); // row bytes
        static constexpr int RP = N * 4 + ((RB % 64) == 0 ? 32 / static_cast<int>(sizeof(S)) : 0);
        // position of node k, equation q in row l
        __device__ static __forceinline__ int pos(int l, int k, int q) { return 4 * k + (q ^ (l & 3)); }
        S         U[EPB][N][RP]; // state; complex-step curved tiles: xi-line part of the volume term after the face phase
        S         F[EPB][N][RP]; // affine meshes: xi-line part of the volume term after the line phase
        S         G[EPB][N][RP]; // affine meshes: eta-line part of the volume term after the line phase
        S         V[CURVED ? EPB : 1][N][CURVED ? N * 4 : 1]; // curved: eta-line part (double: xi + eta part)
        S         Del[EPB][4][N][4];
        double    dL[N], dR[N];
        long long eid[EPB];      // global element of each tile slot, -1 if outside the mesh
    };

    // Metric pair of a curved line task, from the per-node metrics in HBM:
    // xi-line (row ln): (xi_x, xi_y) of nodes (k, ln); eta-line (column ln):
    // (eta_x, eta_y) of nodes (ln, l).
    template <int N>
    __device__ __forceinline__ void line_metrics(const double* __restrict__ vmet, long long e, bool eta, int ln,
                                                 double ma[N], double mb[N])
    {
        constexpr int NN = N * N;
        const double* b  = vmet + (e * NN) * 4 + (eta ? 2 : 0);
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
            const int    node = eta ? m * N + ln : ln * N + m;
            const double2 v   = *reinterpret_cast<const double2*>(b + node * 4);
            ma[m]             = v.x;
            mb[m]             = v.y;
        }
    }

    // Tile origin and shape.  Tiles are numbered strip by strip: the tile
    // columns form eight vertical strips of ceil(ntx / 8) tiles (the last one
    // narrower), each walked row-major, so that the band of tiles one XCD
    // processes (xcd_tile) is one strip and the tile above a tile follows it
    // a strip width later, while its bottom trace is still in that XCD's L2.
    // The last tile of a row or column may hang over the mesh (its outside
    // slots are idle).
    struct TileGeom
    {
        int i0, j0, tx, ty;
    };
    __device__ __forceinline__ TileGeom tile_geom(const Res2DP& P, long long blk, int tx, int ty)
    {
        const int ntx = (P.nx + tx - 1) / tx;
        const int nty = (P.ny + ty - 1) / ty;
        const int sw  = (ntx + 7) / 8;
        const int b   = static_cast<int>(blk);
        const int st  = b / (sw * nty);
        const int w   = min(sw, ntx - st * sw);
        const int r   = b - st * sw * nty;
        const int tj  = r / w;
        return {(st * sw + r - tj * w) * tx, tj * ty, tx, ty};
    }

    // One face-node task of a tile.  Each face of the tile's elements is
    // evaluated once: faces between two elements of the tile take both
    // traces from LDS; faces on the tile rim gather the outer trace from HBM
    // (issued at kernel start) or form it from the boundary condition.
    template <class S>
    struct FaceTask
    {
        bool   live, xdir;
        int    idx;
        int    elA, elB;   // tile slots left/below (A) and right/above (B), -1 if outside the tile
        int    bdir;       // boundary direction if the outer side is a physical boundary, else -1
        double fnx, fny, sc;
        S      Uext[4];    // outer trace gathered from HBM
    };

    template <int N, class IO>
    __device__ __forceinline__ void face_task(const Res2DP& P, const IO& io, const TileGeom& g, int t,
                                              FaceTask<typename IO::S>& f)
    {
        constexpr int NN = N * N;
        const int     XT = (g.tx + 1) * g.ty * N;
        const int     YT = g.tx * (g.ty + 1) * N;
        f.live           = false;
        if (t >= XT + YT) return;
        f.xdir = t < XT;
        f.bdir = -1;
        if (f.xdir)
        {
            // face between slot columns a-1 and a of tile row b
            const int b  = t / ((g.tx + 1) * N);
            const int r  = t - b * (g.tx + 1) * N;
            const int a  = r / N;
            f.idx        = r - a * N;
            const int j  = g.j0 + b;
            const int fi = g.i0 + a; // face index in [0, nx]
            if (j >= P.ny) return;
            const bool inA = a >= 1 && fi - 1 < P.nx;
            const bool inB = a < g.tx && fi < P.nx;
            if (!inA && !inB) return;
            f.live = true;
            f.elA  = inA ? b * g.tx + a - 1 : -1;
            f.elB  = inB ? b * g.tx + a : -1;
            if (P.curved)
            {
                const double* fm = P.fxm + ((static_cast<long long>(j) * (P.nx + 1) + fi) * N + f.idx) * 4;
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
            int ni = -1, nk = 0;
            if (!inA)
            {
                ni = fi - 1;
                nk = N - 1;
                if (ni < 0)
                {
                    if (P.per_x)
                        ni = P.nx - 1;
                    else
                    {
                        ni     = -1;
                        f.bdir = 0;
                    }
                }
            }
            else if (!inB)
            {
                ni = fi;
                nk = 0;
                if (ni >= P.nx)
                {
                    if (P.per_x)
                        ni = 0;
                    else
                    {
                        ni     = -1;
                        f.bdir = 1;
                    }
                }
            }
            if (ni >= 0) io.ld4(((static_cast<long long>(j) * P.nx + ni) * NN + f.idx * N + nk) * 4, f.Uext);
        }
        else
        {
            // face between slot rows bf-1 and bf of tile column a
            const int u  = t - XT;
            const int bf = u / (g.tx * N);
            const int r  = u - bf * g.tx * N;
            const int a  = r / N;
            f.idx        = r - a * N;
            const int i  = g.i0 + a;
            const int fj = g.j0 + bf; // face index in [0, ny]
            if (i >= P.nx) return;
            const bool inA = bf >= 1 && fj - 1 < P.ny;
            const bool inB = bf < g.ty && fj < P.ny;
            if (!inA && !inB) return;
            f.live = true;
            f.elA  = inA ? (bf - 1) * g.tx + a : -1;
            f.elB  = inB ? bf * g.tx + a : -1;
            if (P.curved)
            {
                const double* fm = P.fym + ((static_cast<long long>(fj) * P.nx + i) * N + f.idx) * 4;
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
            int nj = -1, nl = 0;
            if (!inA)
            {
                nj = fj - 1;
                nl = N - 1;
                if (nj < 0)
                {
                    if (P.per_y)
                        nj = P.ny - 1;
                    else
                    {
                        nj     = -1;
                        f.bdir = 3;
                    }
                }
            }
            else if (!inB)
            {
                nj = fj;
                nl = 0;
                if (nj >= P.ny)
                {
                    if (P.per_y)
                        nj = 0;
                    else
                    {
                        nj     = -1;
                        f.bdir = 2;
                    }
                }
            }
            if (nj >= 0) io.ld4(((static_cast<long long>(nj) * P.nx + i) * NN + nl * N + f.idx) * 4, f.Uext);
        }
    }

    // Riemann flux of a face task -> scaled flux jumps of the tile elements on
    // either side (Del[el][face][idx]: faces 0 left, 1 right, 2 bottom, 3 top)
    template <int N, class IO, bool CURVED>
    __device__ __forceinline__ void face_flux(const Res2DP& P, FaceTask<typename IO::S>& f,
                                              TileSmem<N, typename IO::S, CURVED>& sm)
    {
        using S = typename IO::S;
        // trace nodes of side A (left / below) and side B (right / above)
        const int kA = f.xdir ? N - 1 : f.idx, lA = f.xdir ? f.idx : N - 1;
        const int kB = f.xdir ? 0 : f.idx, lB = f.xdir ? f.idx : 0;
        S         UA[4], UB[4], Fs[4];
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
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) UA[q] = f.Uext[q];
            }
        }
        if (f.elB < 0)
        {
            if (f.bdir >= 0)
                bc_exterior(P.bc, f.bdir, UA, f.fnx, f.fny, UB);
            else
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) UB[q] = f.Uext[q];
            }
        }
        riemann2(P.gas, P.riemann, UA, UB, f.fnx, f.fny, Fs);
        const int fA = f.xdir ? 1 : 3, fB = f.xdir ? 0 : 2;
        // own-side normal flux: curved tiles read F, G from LDS; affine tiles
        // recompute them from the trace (the line tasks run concurrently and
        // overwrite F, G in place)
        S FA[4], GA[4], FB[4], GB[4];
        if constexpr (!CURVED)
        {
            if (f.elA >= 0) flux2(P.gas, UA, FA, GA);
            if (f.elB >= 0) flux2(P.gas, UB, FB, GB);
        }
        if (f.elA >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                S fn;
                if constexpr (CURVED)
                    fn = sm.F[f.elA][lA][sm.pos(lA, kA, q)] * f.fnx + sm.G[f.elA][lA][sm.pos(lA, kA, q)] * f.fny;
                else
                    fn = FA[q] * f.fnx + GA[q] * f.fny;
                sm.Del[f.elA][fA][f.idx][q] = f.sc * (Fs[q] - fn);
            }
        }
        if (f.elB >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                S fn;
                if constexpr (CURVED)
                    fn = sm.F[f.elB][lB][sm.pos(lB, kB, q)] * f.fnx + sm.G[f.elB][lB][sm.pos(lB, kB, q)] * f.fny;
                else
                    fn = FB[q] * f.fnx + GB[q] * f.fny;
                sm.Del[f.elB][fB][f.idx][q] = f.sc * (Fs[q] - fn);
            }
        }
    }

    // Even-odd form of an n x n line operator M on the symmetric GLL nodes,
    // whose entries mirror as M[n-1-k][n-1-m] = -M[k][m] (ANTI: the derivative
    // matrix) or +M[k][m] (SYM: the filter).  With h = floor(n/2), the mirrored
    // sums e[m] = x[m] + x[n-1-m] and differences d[m] = x[m] - x[n-1-m]
    // (m < h), and the table halves A[k][m] = (M[k][m] + M[n-1-k][m]) / 2,
    // B[k][m] = (M[k][m] - M[n-1-k][m]) / 2 (k, m < h):
    //   y[k] = a + b,  y[n-1-k] = a - b,
    //   ANTI: a = A d, b = B e + C[k] x[c];  SYM: a = A e + C[k] x[c], b = B d,
    // and for odd n the middle node c = h adds C[k] = M[k][c] (k < h) and
    //   y[c] = R . (ANTI ? d : e) + M[c][c] x[c],  R[m] = M[c][m].
    // Per line: n^2/2 + O(n) operations instead of n^2.  Table of one operator
    // (eo_tab): A [h*h], B [h*h], then for odd n C [h], R [h], M[c][c].
    template <int N>
    __host__ __device__ constexpr int eo_size()
    {
        return 2 * (N / 2) * (N / 2) + ((N % 2) ? 2 * (N / 2) + 1 : 0);
    }
    // the operator's table: which 0 = the derivative matrix, 1 = the filter,
    // 2 = the derivative matrix transposed
    template <int N>
    __device__ __forceinline__ const double* eo_tab(const double* tab, int which)
    {
        return tab + 2 * N * N + 2 * N + which * eo_size<N>();
    }
    // mirrored sums and differences of a line, ordered (for a, for b)
    template <int N, class S, bool ANTI>
    __device__ __forceinline__ void eo_split(const S x[N], S pa[N / 2], S pb[N / 2])
    {
#pragma unroll
        for (int m = 0; m < N / 2; ++m)
        {
            const S e = x[m] + x[N - 1 - m], d = x[m] - x[N - 1 - m];
            pa[m]     = ANTI ? d : e;
            pb[m]     = ANTI ? e : d;
        }
    }
    // outputs k and n-1-k (k < h) from the split line; xc the middle value (odd n)
    template <int N, class S, bool ANTI>
    __device__ __forceinline__ void eo_pair(const double* __restrict__ T, int k, const S pa[N / 2], const S pb[N / 2],
                                            S xc, S& y0, S& y1)
    {
        constexpr int H = N / 2;
        const double* A = T;
        const double* B = T + H * H;
        S             a = A[k * H] * pa[0], b = B[k * H] * pb[0];
#pragma unroll
        for (int m = 1; m < H; ++m)
        {
            a = a + A[k * H + m] * pa[m];
            b = b + B[k * H + m] * pb[m];
        }
        if constexpr (N % 2 == 1)
        {
            const double c = T[2 * H * H + k];
            if constexpr (ANTI)
                b = b + c * xc;
            else
                a = a + c * xc;
        }
        y0 = a + b;
        y1 = a - b;
    }
    // middle output (odd n): R . d (ANTI) or R . e (SYM), both held in pa
    template <int N, class S>
    __device__ __forceinline__ S eo_mid(const double* __restrict__ T, const S pa[N / 2], S xc)
    {
        constexpr int H = N / 2;
        const double* R = T + 2 * H * H + H;
        S             y = T[2 * H * H + 2 * H] * xc;
#pragma unroll
        for (int m = 0; m < H; ++m) y = y + R[m] * pa[m];
        return y;
    }
    // y = M x for a whole line
    template <int N, class S, bool ANTI>
    __device__ __forceinline__ void eo_line(const double* __restrict__ T, const S x[N], S y[N])
    {
        constexpr int H = N / 2;
        S             pa[H > 0 ? H : 1], pb[H > 0 ? H : 1];
        eo_split<N, S, ANTI>(x, pa, pb);
        const S xc = x[H];
#pragma unroll
        for (int k = 0; k < H; ++k) eo_pair<N, S, ANTI>(T, k, pa, pb, xc, y[k], y[N - 1 - k]);
        if constexpr (N % 2 == 1) y[H] = eo_mid<N, S>(T, pa, xc);
    }

    // Volume term of one line of one element (equation q) of the double
    // tiles by the even-odd contraction.  Curved: F and G along the line, the
    // metric pair (ma, mb) of the line's nodes; ACC adds the result into V,
    // else the xi-line result goes to U and the eta-line result to V.
    // Affine: xi-line (row ln) of F or eta-line (column ln) of G in place,
    // scaled by the constant metric.
    template <int N, bool ETA, bool CURVED, bool ACC, class SM>
    __device__ __forceinline__ void volume_line_eo(const Res2DP& P, SM& sm, int le, int ln, int q,
                                                   const double* ma, const double* mb)
    {
        constexpr int H  = N / 2;
        const double* T  = eo_tab<N>(P.tab, 0);
        auto          at = [&](int i, int& l0, int& k0)
        {
            l0 = ETA ? i : ln;
            k0 = ETA ? ln : i;
        };
        if constexpr (CURVED)
        {
            double pf[H], qf[H], pg[H], qg[H], fc = 0.0, gc = 0.0;
#pragma unroll
            for (int m = 0; m < H; ++m)
            {
                int l0, k0, l1, k1;
                at(m, l0, k0);
                at(N - 1 - m, l1, k1);
                const double f0 = sm.F[le][l0][sm.pos(l0, k0, q)], f1 = sm.F[le][l1][sm.pos(l1, k1, q)];
                const double g0 = sm.G[le][l0][sm.pos(l0, k0, q)], g1 = sm.G[le][l1][sm.pos(l1, k1, q)];
                pf[m]           = f0 - f1;
                qf[m]           = f0 + f1;
                pg[m]           = g0 - g1;
                qg[m]           = g0 + g1;
            }
            if constexpr (N % 2 == 1)
            {
                int lc, kc;
                at(H, lc, kc);
                fc = sm.F[le][lc][sm.pos(lc, kc, q)];
                gc = sm.G[le][lc][sm.pos(lc, kc, q)];
            }
            auto put = [&](int i, double v)
            {
                int l0, k0;
                at(i, l0, k0);
                if constexpr (ACC)
                    atomicAdd(&sm.V[le][l0][sm.pos(l0, k0, q)], v);
                else if constexpr (ETA)
                    sm.V[le][l0][sm.pos(l0, k0, q)] = v;
                else
                    sm.U[le][l0][sm.pos(l0, k0, q)] = v;
            };
#pragma unroll
            for (int k = 0; k < H; ++k)
            {
                double f0, f1, g0, g1;
                eo_pair<N, double, true>(T, k, pf, qf, fc, f0, f1);
                eo_pair<N, double, true>(T, k, pg, qg, gc, g0, g1);
                put(k, ma[k] * f0 + mb[k] * g0);
                put(N - 1 - k, ma[N - 1 - k] * f1 + mb[N - 1 - k] * g1);
            }
            if constexpr (N % 2 == 1)
                put(H, ma[H] * eo_mid<N, double>(T, pf, fc) + mb[H] * eo_mid<N, double>(T, pg, gc));
        }
        else
        {
            // the line of F (xi) or G (eta); only this task reads it in this phase
            auto&        X  = ETA ? sm.G : sm.F;
            const double sc = ETA ? P.eta_y : P.xi_x;
            double       x[N], y[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                int l0, k0;
                at(m, l0, k0);
                x[m] = X[le][l0][sm.pos(l0, k0, q)];
            }
            eo_line<N, double, true>(T, x, y);
#pragma unroll
            for (int k = 0; k < N; ++k)
            {
                int l0, k0;
                at(k, l0, k0);
                X[le][l0][sm.pos(l0, k0, q)] = sc * y[k];
            }
        }
    }

    // Volume term along one line of one element for equation q.
    //   xi-line  (row l):    U[l][k] <- xi_x d/dxi F + xi_y d/dxi G   (k = 0..N-1)
    //   eta-line (column k): V[l][k] <- eta_x d/deta F + eta_y d/deta G (l = 0..N-1)
    //   (affine: F[l][k] <- xi_x d/dxi F and G[l][k] <- eta_y d/deta G, in place)
    // The line is held in registers; D comes through the scalar cache, the
    // metric pair (ma, mb) of the line's nodes from registers.  Affine meshes
    // keep only d/dxi F and d/deta G (the cross terms are exact zeros).
    //   With ACC (curved, double) both parts are added into V, zeroed per node
    //   beforehand, by LDS atomic adds: V = fl(xi + eta) whichever lands first,
    //   and U keeps the state, so the line tasks can run beside the face tasks.
    template <int N, class S, bool CURVED, class SM, bool ACC = false>
    __device__ __forceinline__ void volume_line(const Res2DP& P, SM& sm, bool eta, int le,
                                                int ln, int q, const double* ma = nullptr, const double* mb = nullptr)
    {
        const double* D = P.tab;
        if constexpr (std::is_same_v<S, double> && N >= 2)
        {
            // double tiles: even-odd contraction
            if (eta)
                volume_line_eo<N, true, CURVED, ACC>(P, sm, le, ln, q, ma, mb);
            else
                volume_line_eo<N, false, CURVED, ACC>(P, sm, le, ln, q, ma, mb);
            return;
        }
        if (!eta)
        {
            S f[N], g[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                f[m] = sm.F[le][ln][sm.pos(ln, m, q)];
                if constexpr (CURVED) g[m] = sm.G[le][ln][sm.pos(ln, m, q)];
            }
#pragma unroll
            for (int k = 0; k < N; ++k)
            {
                S dF = Sc<S>::c(0.0), dG = Sc<S>::c(0.0);
#pragma unroll
                for (int m = 0; m < N; ++m)
                {
                    dF = dF + D[k * N + m] * f[m];
                    if constexpr (CURVED) dG = dG + D[k * N + m] * g[m];
                }
                // affine: the result replaces row ln of F (equation q), which
                // only this task reads in this phase
                if constexpr (ACC)
                    atomicAdd(&sm.V[le][ln][sm.pos(ln, k, q)], ma[k] * dF + mb[k] * dG);
                else if constexpr (CURVED)
                    sm.U[le][ln][sm.pos(ln, k, q)] = ma[k] * dF + mb[k] * dG;
                else
                    sm.F[le][ln][sm.pos(ln, k, q)] = P.xi_x * dF;
            }
        }
        else
        {
            S f[N], g[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                g[m] = sm.G[le][m][sm.pos(m, ln, q)];
                if constexpr (CURVED) f[m] = sm.F[le][m][sm.pos(m, ln, q)];
            }
#pragma unroll
            for (int l = 0; l < N; ++l)
            {
                S dF = Sc<S>::c(0.0), dG = Sc<S>::c(0.0);
#pragma unroll
                for (int m = 0; m < N; ++m)
                {
                    dG = dG + D[l * N + m] * g[m];
                    if constexpr (CURVED) dF = dF + D[l * N + m] * f[m];
                }
                // affine: the result replaces column ln of G (equation q), which
                // only this task reads in this phase
                if constexpr (ACC)
                    atomicAdd(&sm.V[le][l][sm.pos(l, ln, q)], ma[l] * dF + mb[l] * dG);
                else if constexpr (CURVED)
                    sm.V[le][l][sm.pos(l, ln, q)] = ma[l] * dF + mb[l] * dG;
                else
                    sm.G[le][l][sm.pos(l, ln, q)] = P.eta_y * dG;
            }
        }
    }

    // Spectral filter of the stage value along one line (equation q), the
    // operator from the scalar cache: xi-line (row l) -> F[l][k] = sum_c
    // Fm[k][c] U[l][c]; eta-line (column k) -> G[l][k] = sum_c Fm[l][c] U[c][k].
    template <int N, bool CURVED>
    __device__ __forceinline__ void filter_line(const Res2DP& P, TileSmem<N, double, CURVED>& sm, bool eta, int le,
                                                int ln, int q)
    {
        const double* Fm = P.tab + N * N + 2 * N;
        double        x[N];
#pragma unroll
        for (int c = 0; c < N; ++c) x[c] = eta ? sm.U[le][c][sm.pos(c, ln, q)] : sm.U[le][ln][sm.pos(ln, c, q)];
        if constexpr (N >= 2)
        {
            double y[N];
            eo_line<N, double, false>(eo_tab<N>(P.tab, 1), x, y);
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                if (eta)
                    sm.G[le][r][sm.pos(r, ln, q)] = y[r];
                else
                    sm.F[le][ln][sm.pos(ln, r, q)] = y[r];
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < N; ++r)
        {
            double a = 0.0;
#pragma unroll
            for (int c = 0; c < N; ++c) a += Fm[r * N + c] * x[c];
            if (eta)
                sm.G[le][r][sm.pos(r, ln, q)] = a;
            else
                sm.F[le][ln][sm.pos(ln, r, q)] = a;
        }
    }

    // One tile = a tx x ty block of elements of one mesh, processed by one
    // workgroup (tx * ty <= EPB; the batched sweep may run smaller tiles).
    // With FILT and `filt` (stage outputs only) the element-local spectral
    // filter of the post-step hook is applied in LDS before the store.
    // Phases: (1) node state -> LDS, fluxes F, G -> LDS (outer traces of the
    // rim faces already in flight); (2) face Riemann fluxes -> Del; (3) volume
    // term by xi- and eta-lines; (4) per node: volume + VCJH lift -> output
    // [(5) stage value -> LDS, filter by xi- and eta-lines, store].
    template <int N, class IO, bool CURVED, bool FILT>
    __device__ __forceinline__ void res2d_tile(const Res2DP& P, const IO& io, long long blk, int tx, int ty, bool filt,
                                               TileSmem<N, typename IO::S, CURVED>& sm)
    {
        using S           = typename IO::S;
        constexpr int BS  = Tile2D<N, S>::BS;
        constexpr int NN  = N * N;
        // curved double tiles add the xi- and eta-parts of the volume term into V
        constexpr bool ACC = CURVED && std::is_same_v<S, double>;
        const int     tid = threadIdx.x;
        if (tid < N)
        {
            sm.dL[tid] = P.tab[NN + tid];
            sm.dR[tid] = P.tab[NN + N + tid];
        }
        const TileGeom g    = tile_geom(P, blk, tx, ty);
        const int      el   = tid / NN;
        const int      node = tid - el * NN;
        const int      k    = node % N;
        const int      l    = node / N;
        const int      eb   = el / tx;
        const int      ea   = el - eb * tx;
        const bool     act  = (el < tx * ty) && (g.i0 + ea < P.nx) && (g.j0 + eb < P.ny);
        const long long e   = static_cast<long long>(g.j0 + eb) * P.nx + (g.i0 + ea);
        if (node == 0 && el < tx * ty) sm.eid[el] = act ? e : -1;

        FaceTask<S> f0;
        face_task<N, IO>(P, io, g, tid, f0);

        // Task split of the accumulating (ACC) tiles whose face tasks fit in
        // the first half of the workgroup and whose line tasks fit twice in
        // the second half (n = 8: 96 face tasks, 256 line tasks): waves 0-1
        // evaluate the faces, waves 2-3 the lines, two each, concurrently.
        // Otherwise every thread takes a face task (if any) and then a line task.
        const int  nlt   = tx * ty * N * 4;
        const int  nft   = ((tx + 1) * ty + tx * (ty + 1)) * N;
        const bool split = ACC && BS == 256 && nft <= BS / 2 && 2 * nlt <= BS;
        const int  lt0   = split ? tid - BS / 2 : tid; // first line task of this thread (< 0: none)
        const int  lstep = split ? BS / 2 : BS;
        // curved: metrics of this thread's first line task, in flight with the state
        double ma[CURVED ? N : 1], mb[CURVED ? N : 1];
        if constexpr (CURVED)
        {
            const bool eta = lt0 >= nlt;
            const int  w   = eta ? lt0 - nlt : lt0;
            const int  le  = (w >> 2) / N;
            const int  lb  = le / tx;
            const int  la  = le - lb * tx;
            if (lt0 >= 0 && lt0 < 2 * nlt && g.i0 + la < P.nx && g.j0 + lb < P.ny)
                line_metrics<N>(P.vmet, static_cast<long long>(g.j0 + lb) * P.nx + (g.i0 + la), eta,
                                (w >> 2) - le * N, ma, mb);
        }

        S own[4];
        if (act)
        {
            io.ld4((e * NN + node) * 4, own);
            S F[4], G[4];
            flux2(P.gas, own, F, G);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                sm.U[el][l][sm.pos(l, k, q)] = own[q];
                sm.F[el][l][sm.pos(l, k, q)] = F[q];
                sm.G[el][l][sm.pos(l, k, q)] = G[q];
                if constexpr (ACC) sm.V[el][l][sm.pos(l, k, q)] = 0.0;
            }
        }
        __syncthreads();

        // ---- face tasks: every face of the tile once ----
        if (f0.live) face_flux<N, IO, CURVED>(P, f0, sm);
        for (int t = tid + BS; t < nft; t += BS)
        {
            FaceTask<S> f;
            face_task<N, IO>(P, io, g, t, f);
            if (f.live) face_flux<N, IO, CURVED>(P, f, sm);
        }
        // complex-step curved line tasks write the xi-part over U, which face
        // tasks read; double curved line tasks add into V and affine line tasks
        // work in place on F and G, neither of which face tasks read
        if constexpr (CURVED && !ACC) __syncthreads();

        // ---- volume term by lines (q fastest).  Curved: xi-line tasks, then
        // eta-line tasks (4 contractions per node); affine: one task does
        // xi-line ln of F and eta-line ln of G (2 contractions per node) ----
        if (split)
        {
            // waves 2-3: line tasks lt0 and lt0 + BS/2 (the second one's
            // metrics loaded after the first one: held across it, they cost
            // spills at n = 8)
            if constexpr (ACC)
            {
                const int t1 = lt0 + BS / 2;
                double    mc[N], md[N];
                const bool eta1 = t1 >= nlt;
                const int  w1   = eta1 ? t1 - nlt : t1;
                const int  le1  = (w1 >> 2) / N;
                const bool has1 = lt0 >= 0 && t1 < 2 * nlt && sm.eid[le1] >= 0;
                if (lt0 >= 0 && lt0 < 2 * nlt)
                {
                    const bool eta = lt0 >= nlt;
                    const int  w   = eta ? lt0 - nlt : lt0;
                    const int  le  = (w >> 2) / N;
                    if (sm.eid[le] >= 0)
                        volume_line<N, S, CURVED, decltype(sm), ACC>(P, sm, eta, le, (w >> 2) - le * N, w & 3, ma, mb);
                }
                if (has1) line_metrics<N>(P.vmet, sm.eid[le1], eta1, (w1 >> 2) - le1 * N, mc, md);
                if (has1) volume_line<N, S, CURVED, decltype(sm), ACC>(P, sm, eta1, le1, (w1 >> 2) - le1 * N, w1 & 3, mc, md);
            }
        }
        else
        {
            const int nt = CURVED ? 2 * nlt : nlt;
            for (int t = lt0; t < nt; t += lstep)
            {
                const bool eta = t >= nlt;
                const int  w   = eta ? t - nlt : t;
                const int  q   = w & 3;
                const int  le  = (w >> 2) / N;
                const int  ln  = (w >> 2) - le * N;
                if (sm.eid[le] < 0) continue;
                if constexpr (CURVED)
                {
                    if (t >= lstep) line_metrics<N>(P.vmet, sm.eid[le], eta, ln, ma, mb); // later passes
                    volume_line<N, S, CURVED, decltype(sm), ACC>(P, sm, eta, le, ln, q, ma, mb);
                }
                else
                {
                    volume_line<N, S, CURVED>(P, sm, false, le, ln, q);
                    volume_line<N, S, CURVED>(P, sm, true, le, ln, q);
                }
            }
        }
        // the output's extra operand (stage u0, ifunction udot) of this node,
        // in flight across the barrier and the lift phase
        double z[4];
        if constexpr (IO::pre)
        {
            if (act) io.pre4((e * NN + node) * 4, z);
        }
        __syncthreads();

        S r[4];
        if (act)
        {
            // ACC tiles keep the state in U to the end: reread, not held in registers
            if constexpr (ACC)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) own[q] = sm.U[el][l][sm.pos(l, k, q)];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                if constexpr (ACC)
                    r[q] = sm.V[el][l][sm.pos(l, k, q)];
                else if constexpr (CURVED)
                    r[q] = sm.U[el][l][sm.pos(l, k, q)] + sm.V[el][l][sm.pos(l, k, q)];
                else
                    r[q] = sm.F[el][l][sm.pos(l, k, q)] + sm.G[el][l][sm.pos(l, k, q)];
                r[q] = r[q] + sm.Del[el][0][l][q] * sm.dL[k];
                r[q] = r[q] + sm.Del[el][1][l][q] * sm.dR[k];
                r[q] = r[q] + sm.Del[el][2][k][q] * sm.dL[l];
                r[q] = r[q] + sm.Del[el][3][k][q] * sm.dR[l];
            }
            if (P.rinv)
            {
                S src[4];
                radial_source2(P.gas, own, P.rinv[e * NN + node], src);
#pragma unroll
                for (int q = 0; q < 4; ++q) r[q] = r[q] + src[q];
            }
        }
        if constexpr (FILT && IO::can_filter)
        {
            if (filt)
            {
                // stage value -> LDS (each node overwrites only its own slot of
                // U, which only this thread reads in this phase; F and G are
                // free after the previous barrier), then out = 1/2 (F_x v + F_y v)
                if (act)
                {
                    double o[4];
                    io.val4p(own, r, z, o);
#pragma unroll
                    for (int q = 0; q < 4; ++q) sm.U[el][l][sm.pos(l, k, q)] = o[q];
                }
                __syncthreads();
                for (int t = tid; t < 2 * nlt; t += BS)
                {
                    const bool eta = t >= nlt;
                    const int  w   = eta ? t - nlt : t;
                    const int  le  = (w >> 2) / N;
                    if (sm.eid[le] < 0) continue;
                    filter_line<N, CURVED>(P, sm, eta, le, (w >> 2) - le * N, w & 3);
                }
                __syncthreads();
                if (act)
                {
                    double o[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) o[q] = 0.5 * (sm.F[el][l][sm.pos(l, k, q)] + sm.G[el][l][sm.pos(l, k, q)]);
                    io.put4((e * NN + node) * 4, o);
                }
                return;
            }
        }
        if (act)
        {
            if constexpr (IO::pre)
                io.st4p((e * NN + node) * 4, own, r, z);
            else
                io.st4((e * NN + node) * 4, own, r);
        }
    }

    // Waves per SIMD that the LDS of one curved tile allows (160 KB per CU,
    // four waves per workgroup), at most 4, for n <= 8: the register budget
    // of the double residual and stage kernels is held to that occupancy, so
    // the filter's line tasks (compiled into every stage launch) cannot cost
    // a wave (n = 8: 130 -> 128 VGPRs, 3 -> 4 waves).  Affine tiles and n = 9
    // stay unconstrained (1): a 128-register cap spills there.
    template <int N, bool CURVED>
    constexpr int stage_waves()
    {
        constexpr int wg = static_cast<int>((160 * 1024) / sizeof(TileSmem<N, double, CURVED>));
        if (!CURVED || N > 8) return 1;
        return wg < 1 ? 1 : (wg > 4 ? 4 : wg);
    }
    // ------------------------------------------------------------------
    // 2D residual: physics_impl_2D.hpp:267-357 (volume) + :689-899 (faces)
    // ------------------------------------------------------------------
    template <int N, class IO, bool CURVED>
    __global__ __launch_bounds__(BlockSize<typename IO::S>::value)
    __attribute__((amdgpu_waves_per_eu(std::is_same_v<typename IO::S, double> ? stage_waves<N, CURVED>() : 1)))
    void k_res2d(const Res2DP P, const IO io)
    {
        using T = Tile2D<N, typename IO::S>;
        __shared__ TileSmem<N, typename IO::S, CURVED> sm;
        res2d_tile<N, IO, CURVED, false>(P, io, xcd_tile(blockIdx.x, gridDim.x), T::TX, T::TY, false, sm);
    }
    // SSP-RK stage fused with the residual; with FILT and `filt` (the last
    // stage of a step when the spectral filter is on) the post-step filter is
    // applied in LDS before the store.  Curved meshes run every stage of a
    // step through the FILT instantiation (one kernel, a uniform branch);
    // affine meshes compile the filter only into the stage that uses it (the
    // combined affine kernel spills 36 bytes at n = 6, 8).
    template <int N, bool CURVED, bool FILT>
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stage_waves<N, CURVED>())))
    void k_res2d_stage(const Res2DP P, const IOStage io, int filt)
    {
        using T = Tile2D<N, double>;
        __shared__ TileSmem<N, double, CURVED> sm;
        // affine FILT kernels always filter (the unfiltered store is dead code there)
        const bool f = CURVED ? filt != 0 : FILT;
        res2d_tile<N, IOStage, CURVED, FILT>(P, io, xcd_tile(blockIdx.x, gridDim.x), T::TX, T::TY, f, sm);
    }

    // ------------------------------------------------------------------
    // 2D forward tangent out = sigma v + R'(u) v, the directional derivative of
    // the residual above (the reference forms J v by the complex step,
    // petsc_solver_impl.hpp:289-356).  The residual is linear in the fluxes
    // except at the node fluxes and the face Riemann fluxes, so only those run
    // in forward-mode pairs (value, derivative); everything after them, the
    // derivative-matrix contraction, the lifts and the output, is the
    // residual's own linear code on the derivative part alone.  LDS: U primal
    // state, V tangent state, F / G the tangent fluxes F'(u) v, G'(u) v; the
    // same 256-thread tiles, face tasks and line tasks as the residual.
    // ------------------------------------------------------------------
    // LDS of one tangent tile: the primal and tangent traces of each element
    // face (Tr[el][face][idx]: u then v of the four equations; faces 0 left,
    // 1 right, 2 bottom, 3 top, as Del), which is all the face tasks read of
    // the states, tangent fluxes F'(u) v, G'(u) v, and (curved) V, the volume
    // term, both line parts added into it by LDS atomics (zeroed per node
    // first); the residual tile's padding and equation order.  Keeping the
    // traces instead of both whole states holds the curved n = 8 tile to
    // 39.1 KB (four tiles per CU).
    template <int N, bool CURVED>
    struct TanSmem
    {
        static constexpr int NN = N * N, EPB = Tile2D<N, double>::EPB;
        static constexpr int RP = TileSmem<N, double, true>::RP;
        __device__ static __forceinline__ int pos(int l, int k, int q) { return 4 * k + (q ^ (l & 3)); }
        double    Tr[EPB][4][N][8];
        double    F[EPB][N][RP];
        double    G[EPB][N][RP];
        double    V[CURVED ? EPB : 1][N][CURVED ? N * 4 : 1];
        double    Del[EPB][4][N][4];
        double    dL[N], dR[N];
        long long eid[EPB];
    };

    template <int N, bool CURVED>
    __device__ __forceinline__ void face_flux_tan(const Res2DP& P, FaceTask<Dual>& f, TanSmem<N, CURVED>& sm)
    {
        const int kA = f.xdir ? N - 1 : f.idx, lA = f.xdir ? f.idx : N - 1;
        const int kB = f.xdir ? 0 : f.idx, lB = f.xdir ? f.idx : 0;
        const int fA = f.xdir ? 1 : 3, fB = f.xdir ? 0 : 2;
        Dual      UA[4], UB[4], Fs[4];
        if (f.elA >= 0)
        {
            const double* t = sm.Tr[f.elA][fA][f.idx];
#pragma unroll
            for (int q = 0; q < 4; ++q) UA[q] = {t[q], t[4 + q]};
        }
        if (f.elB >= 0)
        {
            const double* t = sm.Tr[f.elB][fB][f.idx];
#pragma unroll
            for (int q = 0; q < 4; ++q) UB[q] = {t[q], t[4 + q]};
        }
        if (f.elA < 0)
        {
            if (f.bdir >= 0)
                bc_exterior(P.bc, f.bdir, UB, f.fnx, f.fny, UA);
            else
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) UA[q] = f.Uext[q];
            }
        }
        if (f.elB < 0)
        {
            if (f.bdir >= 0)
                bc_exterior(P.bc, f.bdir, UA, f.fnx, f.fny, UB);
            else
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) UB[q] = f.Uext[q];
            }
        }
        riemann2(P.gas, P.riemann, UA, UB, f.fnx, f.fny, Fs);
        // own-side normal flux tangent: curved tiles read F'v, G'v from LDS;
        // affine tiles find it in their own Del slot, put there by the trace
        // node (the line tasks overwrite F, G in place meanwhile)
        if (f.elA >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                double fn;
                if constexpr (CURVED)
                    fn = sm.F[f.elA][lA][sm.pos(lA, kA, q)] * f.fnx + sm.G[f.elA][lA][sm.pos(lA, kA, q)] * f.fny;
                else
                    fn = sm.Del[f.elA][fA][f.idx][q];
                sm.Del[f.elA][fA][f.idx][q] = f.sc * (Fs[q].d - fn);
            }
        }
        if (f.elB >= 0)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                double fn;
                if constexpr (CURVED)
                    fn = sm.F[f.elB][lB][sm.pos(lB, kB, q)] * f.fnx + sm.G[f.elB][lB][sm.pos(lB, kB, q)] * f.fny;
                else
                    fn = sm.Del[f.elB][fB][f.idx][q];
                sm.Del[f.elB][fB][f.idx][q] = f.sc * (Fs[q].d - fn);
            }
        }
    }

    template <int N, bool CURVED>
    __device__ __forceinline__ void tan2d_tile(const Res2DP& P, const IOTan& io, long long blk, int tx, int ty,
                                               TanSmem<N, CURVED>& sm)
    {
        constexpr int BS  = 256;
        constexpr int NN  = N * N;
        const int     tid = threadIdx.x;
        if (tid < N)
        {
            sm.dL[tid] = P.tab[NN + tid];
            sm.dR[tid] = P.tab[NN + N + tid];
        }
        const TileGeom g    = tile_geom(P, blk, tx, ty);
        const int      el   = tid / NN;
        const int      node = tid - el * NN;
        const int      k    = node % N;
        const int      l    = node / N;
        const int      eb   = el / tx;
        const int      ea   = el - eb * tx;
        const bool     act  = (el < tx * ty) && (g.i0 + ea < P.nx) && (g.j0 + eb < P.ny);
        const long long e   = static_cast<long long>(g.j0 + eb) * P.nx + (g.i0 + ea);
        if (node == 0 && el < tx * ty) sm.eid[el] = act ? e : -1;

        FaceTask<Dual> f0;
        face_task<N, IOTan>(P, io, g, tid, f0);

        // curved: the line parts are added into V (no barrier between face and
        // line tasks), and where the face tasks fit in waves 0-1 and the line
        // tasks twice in waves 2-3 (n = 8) the two run on separate waves, as in
        // res2d_tile.  The line metrics are loaded by the line tasks (held
        // through the face phase they cost a wave per SIMD).
        const int  nlt   = tx * ty * N * 4;
        const int  nft   = ((tx + 1) * ty + tx * (ty + 1)) * N;
        const bool split = CURVED && nft <= BS / 2 && 2 * nlt <= BS;
        double     vt[4]; // own tangent state
        if (act)
        {
            Dual own[4], F[4], G[4];
            io.ld4((e * NN + node) * 4, own);
            flux2(P.gas, own, F, G);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                vt[q]                        = own[q].d;
                sm.F[el][l][sm.pos(l, k, q)] = F[q].d;
                sm.G[el][l][sm.pos(l, k, q)] = G[q].d;
                if constexpr (CURVED) sm.V[el][l][sm.pos(l, k, q)] = 0.0;
            }
            // the node's traces (corner nodes lie on two faces)
            auto trace = [&](int fc, int idx)
            {
                double* t = sm.Tr[el][fc][idx];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    t[q]     = own[q].v;
                    t[4 + q] = own[q].d;
                }
            };
            if (k == 0) trace(0, l);
            if (k == N - 1) trace(1, l);
            if (l == 0) trace(2, k);
            if (l == N - 1) trace(3, k);
            if constexpr (!CURVED)
            {
                // affine faces have normals (1, 0) and (0, 1): the trace nodes
                // leave F'v (left, right) or G'v (bottom, top) in the Del slot
                // that the face task replaces with the scaled jump
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    if (k == 0) sm.Del[el][0][l][q] = F[q].d;
                    if (k == N - 1) sm.Del[el][1][l][q] = F[q].d;
                    if (l == 0) sm.Del[el][2][k][q] = G[q].d;
                    if (l == N - 1) sm.Del[el][3][k][q] = G[q].d;
                }
            }
        }
        __syncthreads();

        if (f0.live) face_flux_tan<N, CURVED>(P, f0, sm);
        for (int t = tid + BS; t < nft; t += BS)
        {
            FaceTask<Dual> f;
            face_task<N, IOTan>(P, io, g, t, f);
            if (f.live) face_flux_tan<N, CURVED>(P, f, sm);
        }

        {
            const int nt    = CURVED ? 2 * nlt : nlt;
            const int lt0   = split ? tid - BS / 2 : tid;
            const int lstep = split ? BS / 2 : BS;
            for (int t = lt0; t < nt; t += lstep)
            {
                if (t < 0) break;
                const bool eta = t >= nlt;
                const int  w   = eta ? t - nlt : t;
                const int  q   = w & 3;
                const int  le  = (w >> 2) / N;
                const int  ln  = (w >> 2) - le * N;
                if (sm.eid[le] < 0) continue;
                if constexpr (CURVED)
                {
                    double ma[N], mb[N];
                    line_metrics<N>(P.vmet, sm.eid[le], eta, ln, ma, mb);
                    volume_line<N, double, true, TanSmem<N, CURVED>, true>(P, sm, eta, le, ln, q, ma, mb);
                }
                else
                {
                    volume_line<N, double, false>(P, sm, false, le, ln, q);
                    volume_line<N, double, false>(P, sm, true, le, ln, q);
                }
            }
        }
        __syncthreads();

        if (!act) return;
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            double r;
            if constexpr (CURVED)
                r = sm.V[el][l][sm.pos(l, k, q)];
            else
                r = sm.F[el][l][sm.pos(l, k, q)] + sm.G[el][l][sm.pos(l, k, q)];
            r    = r + sm.Del[el][0][l][q] * sm.dL[k];
            r    = r + sm.Del[el][1][l][q] * sm.dR[k];
            r    = r + sm.Del[el][2][k][q] * sm.dL[l];
            r    = r + sm.Del[el][3][k][q] * sm.dR[l];
            o[q] = r;
        }
        if (P.rinv)
        {
            Dual own[4], src[4];
            io.ld4((e * NN + node) * 4, own);
            radial_source2(P.gas, own, P.rinv[e * NN + node], src);
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = o[q] + src[q].d;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = io.sigma * vt[q] + o[q];
        st4d(io.out + (e * NN + node) * 4, o);
    }
    template <int N, bool CURVED>
    __global__ __launch_bounds__(256) void k_tan2d(const Res2DP P, const IOTan io)
    {
        using T = Tile2D<N, double>;
        __shared__ TanSmem<N, CURVED> sm;
        tan2d_tile<N, CURVED>(P, io, xcd_tile(blockIdx.x, gridDim.x), T::TX, T::TY, sm);
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
        const double* rinv; // [ne][n] 1/radius of radial meshes, else nullptr
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
#pragma unroll
        for (int jn = 0; jn < N; ++jn)
        {
            S nd[3] = {u[0][jn], u[1][jn], u[2][jn]}, F[3];
            flux1(P.gas, nd, F);
#pragma unroll
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
#pragma unroll
            for (int in = 0; in < N; ++in)
                res[q][in] = res[q][in] + P.inv_jac * (dl * dL[in] + dr * dR[in]);
        }
        // radial source, added to the interior elements only (the reference's
        // boundary elements are evaluated before its source loop,
        // physics_impl_1D.hpp:29-94)
        if (P.rinv && (P.periodic || (e > 0 && e < P.ne - 1)))
        {
#pragma unroll
            for (int m = 0; m < N; ++m)
            {
                S nd[3] = {u[0][m], u[1][m], u[2][m]}, src[3];
                radial_source1(P.gas, nd, P.rinv[static_cast<long long>(e) * N + m], src);
#pragma unroll
                for (int q = 0; q < 3; ++q) res[q][m] = res[q][m] + src[q];
            }
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

// frsdg_color.hpp — the preconditioning matrix of the reference's implicit
// path and its incomplete factorisation.
//
// The reference hands PETSc a sparse matrix P that PETSc fills by coloured
// finite differences of the implicit function (DMCreateMatrix +
// DMCreateColoring(IS_COLORING_GLOBAL) + MatFDColoringApply:
// petsc_solver_impl.hpp:72-95, 126-160, 221-237) and, with the default
// -pc_type of a one-process run, factors it by ILU(0) in the natural
// ordering.  The DMDA has one grid point per element with dof = the element's
// DOFs and a box stencil of width 1 (petsc_solver_impl.hpp:496-516), so P is
// block sparse: block row e holds the (nd x nd) blocks dF_e/du_m of the 3x3
// (2D) or 3 (1D) neighbouring elements m, including blocks that the residual
// leaves zero (the diagonal neighbours in 2D).
//
// Storage: P[e][slot] column-major (nd x nd), slot = (dj+1)*3 + (di+1) in 2D,
// di+1 in 1D; slots without a neighbour (non-periodic boundaries) stay zero.
// The colouring is PETSc's for a box stencil of width 1: colour of DOF k of
// element (i, j) = k + nd * (i mod 3 + 3 (j mod 3)) (1D: k + nd (i mod 3)), so
// elements of one colour class are three apart and each block row meets at
// most one of them; PETSc requires periodic directions to hold a multiple of
// 3 elements, and so does this code.
//
// The factorisation is block ILU(0) on that block pattern, which equals
// PETSc's point ILU(0) on the AIJ pattern (every entry of a stored block is
// in the pattern): for each block row I in natural order,
//   L_IK = A_IK U_KK^-1  and  A_IJ -= L_IK U_KJ  (J > K in both patterns)
// for its lower neighbours K in increasing order, then A_II = L_II U_II
// (unpivoted), then U_IJ = L_II^-1 A_IJ for its upper neighbours J.  Rows
// whose lower neighbours are all factored form a level (i + 2j for a
// non-periodic 2D mesh: nx + 2 ny - 2 levels).  A level runs as a short
// sequence of launches whose workgroups are 48 x 48 tiles of the block
// products (the triangular solves use the stored inverses of the factored
// diagonal blocks), one workgroup per row for the diagonal LU and one per
// inverse.  A last, fully parallel pass scales every off-diagonal block by
// its row's diagonal inverse (k_ilu_scale), which makes the rows of both
// sweeps of the apply independent.
#pragma once
#include "../../include/frsdg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace frsdg
{
    constexpr int CMAXN = 8;  // neighbours below / above a block row (3x3 box minus the centre)
    constexpr int CMAXP = 9;  // update pairs per lower neighbour

    struct ColorPlan
    {
        int       S    = 0; // slots per block row (9 in 2D, 3 in 1D)
        int       nd   = 0; // DOFs per element
        int       ncls = 0; // colour classes of elements (9 or 3)
        long long E    = 0;
        std::vector<int> lvl_ptr;   // level offsets into lvl (host)
        std::vector<int> lvl_nlow;  // most lower neighbours of a row, per level (host)
        std::vector<int> lvl_nupp;  // most upper neighbours of a row, per level (host)
        std::vector<int> lvl_npair; // [level][q]: most update pairs of a row's q-th lower neighbour (host)
        std::vector<int> cls_count; // elements per colour class (host)
        int* nbr  = nullptr; // [E][S] neighbour element per slot, -1 none
        int* cls  = nullptr; // [E] colour class
        int* cslot = nullptr; // [ncls][E] slot of row e whose neighbour has colour class c, -1 none
        int* low  = nullptr; // [E][CMAXN] lower slots (neighbour index < e) in increasing neighbour order
        int* upd  = nullptr; // [E][CMAXN][2 CMAXP] (slot of J in row I, slot of J in row K) per lower slot
        int* upp  = nullptr; // [E][CMAXN] upper slots
        int* lvl  = nullptr; // [E] elements grouped by level
        int* lpos = nullptr; // [E][2][CMAXN] per level position: lower (slot, element) pairs, slot -1 ends
        int* upos = nullptr; // [E][2][CMAXN] per level position: upper (slot, element) pairs
        int* info = nullptr; // zero-pivot flag
        int* df   = nullptr; // dependence-ordered sweeps: [0] [1] ticket counters (forward, backward),
                             // [2] wait-limit flag, [3, 3 + 2E) chunks stored per row (forward, backward)
        double* ydf = nullptr; // their forward / backward results, rows padded to df_stride(nd)
        double* zdf = nullptr;
        double* P    = nullptr; // E*S*nd*nd for the Newton preconditioner (allocated on first use)
        double* Dinv = nullptr; // E*2*nd*nd: L_II^-1, U_II^-1 (with P)
        float*  P32    = nullptr; // single-precision copies of P and Dinv after the factorisation
        float*  Dinv32 = nullptr; // (Newton option pc_fp32, allocated on first use)
        double* y    = nullptr; // forward-sweep vector
        double* v    = nullptr; // perturbed state / unit colour vector
        double* f0   = nullptr; // F at the base state
        double* f1   = nullptr; // F at the perturbed state
        double* t    = nullptr; // transposed sweeps: U_JJ^T y_J, then L_KK^T z_K (allocated on first use)
        double* T    = nullptr; // factorisation scratch: tcap blocks
        long long tcap = 0;
    };

    // Neighbour tables, colour classes and the ILU(0) level schedule.
    // Returns FRSDG_ERR_UNSUPPORTED where PETSc's DMDA colouring refuses the
    // mesh (a periodic direction whose element count is not a multiple of 3).
    inline int color_plan_tables(int dims, int nx, int ny, int per_x, int per_y, int nd, ColorPlan& cp,
                                 std::vector<int>& h_nbr, std::vector<int>& h_cls, std::vector<int>& h_low,
                                 std::vector<int>& h_upd, std::vector<int>& h_upp, std::vector<int>& h_lvl,
                                 std::vector<int>& h_lpos, std::vector<int>& h_upos)
    {
        if (dims == 1) ny = 1, per_y = 0;
        if ((per_x && nx % 3) || (per_y && ny % 3)) return FRSDG_ERR_UNSUPPORTED;
        const int S  = dims == 2 ? 9 : 3;
        const long long E = static_cast<long long>(nx) * ny;
        cp.S = S;
        cp.nd = nd;
        cp.E = E;
        cp.ncls = dims == 2 ? 9 : 3;
        h_nbr.assign(E * S, -1);
        h_cls.assign(E, 0);
        cp.cls_count.assign(cp.ncls, 0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
            {
                const long long e = static_cast<long long>(j) * nx + i;
                h_cls[e] = dims == 2 ? (i % 3) + 3 * (j % 3) : i % 3;
                cp.cls_count[h_cls[e]]++;
                for (int dj = (dims == 2 ? -1 : 0); dj <= (dims == 2 ? 1 : 0); ++dj)
                    for (int di = -1; di <= 1; ++di)
                    {
                        int ii = i + di, jj = j + dj;
                        if (ii < 0 || ii >= nx)
                        {
                            if (!per_x) continue;
                            ii = (ii + nx) % nx;
                        }
                        if (jj < 0 || jj >= ny)
                        {
                            if (!per_y) continue;
                            jj = (jj + ny) % ny;
                        }
                        const int slot = dims == 2 ? (dj + 1) * 3 + (di + 1) : di + 1;
                        h_nbr[e * S + slot] = jj * nx + ii;
                    }
            }
        h_low.assign(E * CMAXN, -1);
        h_upp.assign(E * CMAXN, -1);
        h_upd.assign(E * CMAXN * 2 * CMAXP, -1);
        std::vector<int> level(E, 0);
        int nlev = 0;
        for (long long e = 0; e < E; ++e)
        {
            std::vector<std::pair<int, int>> lo, up; // (neighbour, slot)
            for (int s = 0; s < S; ++s)
            {
                const int m = h_nbr[e * S + s];
                if (m < 0 || m == e) continue;
                (m < e ? lo : up).push_back({m, s});
            }
            std::sort(lo.begin(), lo.end());
            int lev = 0;
            for (size_t q = 0; q < lo.size(); ++q)
            {
                const int K = lo[q].first;
                h_low[e * CMAXN + q] = lo[q].second;
                lev = std::max(lev, level[K] + 1);
                int np = 0;
                for (int sj = 0; sj < S; ++sj)
                {
                    const int J = h_nbr[e * S + sj];
                    if (J < 0 || J <= K) continue;
                    for (int tk = 0; tk < S; ++tk)
                        if (h_nbr[static_cast<long long>(K) * S + tk] == J)
                        {
                            const long long at = ((e * CMAXN + q) * CMAXP + np) * 2;
                            h_upd[at]     = sj;
                            h_upd[at + 1] = tk;
                            ++np;
                            break;
                        }
                }
            }
            for (size_t q = 0; q < up.size(); ++q) h_upp[e * CMAXN + q] = up[q].second;
            level[e] = lev;
            nlev = std::max(nlev, lev + 1);
        }
        cp.lvl_ptr.assign(nlev + 1, 0);
        for (long long e = 0; e < E; ++e) cp.lvl_ptr[level[e] + 1]++;
        for (int l = 0; l < nlev; ++l) cp.lvl_ptr[l + 1] += cp.lvl_ptr[l];
        h_lvl.assign(E, 0);
        std::vector<int> fill(cp.lvl_ptr.begin(), cp.lvl_ptr.end() - 1);
        for (long long e = 0; e < E; ++e) h_lvl[fill[level[e]]++] = static_cast<int>(e);
        cp.lvl_nlow.assign(nlev, 0);
        cp.lvl_nupp.assign(nlev, 0);
        cp.lvl_npair.assign(static_cast<size_t>(nlev) * CMAXN, 0);
        for (long long e = 0; e < E; ++e)
        {
            int nl = 0, nu = 0;
            while (nl < CMAXN && h_low[e * CMAXN + nl] >= 0) ++nl;
            while (nu < CMAXN && h_upp[e * CMAXN + nu] >= 0) ++nu;
            cp.lvl_nlow[level[e]] = std::max(cp.lvl_nlow[level[e]], nl);
            cp.lvl_nupp[level[e]] = std::max(cp.lvl_nupp[level[e]], nu);
            for (int q = 0; q < nl; ++q)
            {
                int np = 0;
                while (np < CMAXP && h_upd[((e * CMAXN + q) * CMAXP + np) * 2] >= 0) ++np;
                int& m = cp.lvl_npair[static_cast<size_t>(level[e]) * CMAXN + q];
                m      = std::max(m, np);
            }
        }
        // the sweeps read a row's neighbours by level position, in one round trip
        h_lpos.assign(E * 2 * CMAXN, -1);
        h_upos.assign(E * 2 * CMAXN, -1);
        for (long long pos = 0; pos < E; ++pos)
        {
            const long long e = h_lvl[pos];
            for (int q = 0; q < CMAXN; ++q)
            {
                const int sl = h_low[e * CMAXN + q], su = h_upp[e * CMAXN + q];
                h_lpos[(pos * 2) * CMAXN + q]     = sl;
                h_lpos[(pos * 2 + 1) * CMAXN + q] = sl >= 0 ? h_nbr[e * S + sl] : 0;
                h_upos[(pos * 2) * CMAXN + q]     = su;
                h_upos[(pos * 2 + 1) * CMAXN + q] = su >= 0 ? h_nbr[e * S + su] : 0;
            }
        }
        return FRSDG_OK;
    }

    // ---- coloured assembly ----------------------------------------------------

    // mode 0: w3 = x + h at the colour's DOFs (h = sqrt(1 + ||x||) eps, PETSc's
    // default "wp" differencing); mode 1: per-column h = eps * x_col with |x_col|
    // raised to umin ("ds"); mode 2: unit vector of the colour (w3 starts zero).
    // restore puts back x (modes 0, 1) or zero (mode 2).
    static __global__ __launch_bounds__(256) void k_color_perturb(long long E, int nd, const int* __restrict__ cls,
                                                                  int cc, int kdof, const double* __restrict__ x,
                                                                  double* __restrict__ w3, double h_wp, double eps,
                                                                  double umin, int mode, int restore)
    {
        const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (e >= E || cls[e] != cc) return;
        const long long col = e * nd + kdof;
        if (mode == 2)
        {
            w3[col] = restore ? 0.0 : 1.0;
            return;
        }
        if (restore)
        {
            w3[col] = x[col];
            return;
        }
        if (mode == 0)
            w3[col] = x[col] + h_wp;
        else
        {
            double dx = x[col];
            if (::fabs(dx) < umin) dx = dx >= 0.0 ? umin : -umin;
            dx *= eps;
            const double vscale = 1.0 / dx;
            w3[col] = x[col] + 1.0 / vscale;
        }
    }

    // Column kdof of the blocks whose neighbour is in colour class cc:
    // (F(w3) - F(x)) / h (modes 0, 1) or the product J e_colour (mode 2).
    static __global__ __launch_bounds__(256) void k_color_store(long long E, int S, int nd, const int* __restrict__ nbr,
                                                                const int* __restrict__ cslot, int cc, int kdof,
                                                                const double* __restrict__ f1,
                                                                const double* __restrict__ f0,
                                                                const double* __restrict__ x, double dxinv_wp,
                                                                double eps, double umin, int mode,
                                                                double* __restrict__ P, double* __restrict__ w3)
    {
        const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (idx >= E * nd) return;
        const long long R = idx / nd;
        const int r = static_cast<int>(idx - R * nd);
        // the one neighbour of this block row in colour class cc (each block row
        // meets at most one element of a class)
        const int s = cslot[cc * E + R];
        if (s < 0) return;
        const int m = nbr[R * S + s];
        double val;
        if (mode == 2)
            val = f1[idx];
        else
        {
            const double dy = f1[idx] - f0[idx];
            if (mode == 0)
                val = dy * dxinv_wp;
            else
            {
                double dx = x[static_cast<long long>(m) * nd + kdof];
                if (::fabs(dx) < umin) dx = dx >= 0.0 ? umin : -umin;
                dx *= eps;
                val = dy * (1.0 / dx);
            }
        }
        P[((R * S + s) * nd + kdof) * static_cast<long long>(nd) + r] = val;
        // rows of the class itself put back the colour's entry of the perturbed
        // copy (x, or zero for the unit colour vectors): F(w3) has been read
        if (s == S / 2 && r == kdof) w3[idx] = mode == 2 ? 0.0 : x[idx];
    }

    // ---- block ILU(0) -----------------------------------------------------------

    constexpr int GT = 144; // largest block factored in a register tile (256 threads x 9 x 9)

    // Register tiles of the diagonal blocks (nd <= 144): 256 threads, thread
    // (tx, ty) = (tid & 15, tid >> 4) holding x[i][j] = entry (tx + 16 i,
    // ty + 16 j).  The elimination steps k of one 16-row block KB = k / 16 are
    // instantiated with KB known at compile time, so a step touches only the
    // live part of the tile (rows and columns of blocks >= KB for the LU) with
    // fixed register indices; the pivot row and column go through LDS, double
    // buffered by step (one barrier per step: a thread passing step k+1's
    // barrier has finished reading step k's buffers).

    // A_II = L_II U_II (unpivoted, right-looking), steps of block KB
    template <int KB>
    __device__ __forceinline__ void lu_block_steps(int nd, double (&x)[9][9], double* __restrict__ bufA,
                                                   double* __restrict__ bufB, int* __restrict__ info, int tx, int ty)
    {
#pragma unroll 1
        for (int kt = 0; kt < 16; ++kt)
        {
            const int k = 16 * KB + kt;
            if (k >= nd) return;
            double* rowk = bufA + (k & 1) * GT;
            double* colk = bufB + (k & 1) * GT;
            if (tx == kt)
#pragma unroll
                for (int j = KB; j < 9; ++j) rowk[ty + 16 * j] = x[KB][j];
            if (ty == kt)
#pragma unroll
                for (int i = KB; i < 9; ++i) colk[tx + 16 * i] = x[i][KB];
            __syncthreads();
            const double piv = rowk[k];
            if (threadIdx.x == 0 && ::fabs(piv) <= 1e-12) info[0] = 1; // PETSc's default zero-pivot tolerance
            const double rpiv = 1.0 / piv;
            // rows at or above the pivot get a zero multiplier and columns at or
            // left of it a zero pivot-row entry (selects); blocks below KB are done
            double w[9], l[9];
#pragma unroll
            for (int j = KB; j < 9; ++j)
            {
                const double v = rowk[ty + 16 * j];
                w[j]           = (j > KB || ty > kt) ? v : 0.0;
            }
#pragma unroll
            for (int i = KB; i < 9; ++i)
            {
                const double v = colk[tx + 16 * i];
                l[i]           = ((i > KB || tx > kt) && tx + 16 * i < nd) ? v * rpiv : 0.0;
            }
#pragma unroll
            for (int i = KB; i < 9; ++i)
#pragma unroll
                for (int j = KB; j < 9; ++j) x[i][j] = ::fma(-l[i], w[j], x[i][j]);
            // the pivot column below the diagonal becomes the multipliers
            if (ty == kt)
#pragma unroll
                for (int i = KB; i < 9; ++i) x[i][KB] = (i > KB || tx > kt) ? l[i] : x[i][KB];
        }
    }
    __device__ inline void diag_lu_reg(int nd, double* __restrict__ D, double* __restrict__ bufA,
                                       double* __restrict__ bufB, int* __restrict__ info)
    {
        const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
        double x[9][9];
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j)
            {
                const int r = tx + 16 * i, c = ty + 16 * j;
                x[i][j] = (r < nd && c < nd) ? D[r + static_cast<long long>(c) * nd] : 0.0;
            }
        lu_block_steps<0>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<1>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<2>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<3>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<4>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<5>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<6>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<7>(nd, x, bufA, bufB, info, tx, ty);
        lu_block_steps<8>(nd, x, bufA, bufB, info, tx, ty);
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j)
            {
                const int r = tx + 16 * i, c = ty + 16 * j;
                if (r < nd && c < nd) D[r + static_cast<long long>(c) * nd] = x[i][j];
            }
        __syncthreads();
    }

    // L_II^-1 of the factored block D (unit lower part): X = I; for k
    // ascending, rows r > k: X_r -= L_rk X_k.  Row k of X is zero right of
    // column k, so a step of block KB updates rows of blocks >= KB and columns
    // of blocks <= KB; the column of L is read from D one step ahead.
    template <int KB>
    __device__ __forceinline__ void linv_block_steps(int nd, const double* __restrict__ D, double (&x)[9][9],
                                                     double& nxt, double* __restrict__ bufA,
                                                     double* __restrict__ bufB, int tx, int ty, int tid)
    {
#pragma unroll 1
        for (int kt = 0; kt < 16; ++kt)
        {
            const int k = 16 * KB + kt;
            if (k >= nd - 1) return;
            double* rowk = bufA + (k & 1) * GT;
            double* colk = bufB + (k & 1) * GT;
            if (tx == kt)
#pragma unroll
                for (int j = 0; j <= KB; ++j) rowk[ty + 16 * j] = x[KB][j];
            if (tid > k && tid < nd) colk[tid] = nxt;
            if (tid < nd) nxt = D[tid + static_cast<long long>(k + 1) * nd];
            __syncthreads();
            double w[9], l[9];
#pragma unroll
            for (int j = 0; j <= KB; ++j) w[j] = rowk[ty + 16 * j];
#pragma unroll
            for (int i = KB; i < 9; ++i)
            {
                const double v = colk[tx + 16 * i];
                l[i]           = ((i > KB || tx > kt) && tx + 16 * i < nd) ? v : 0.0;
            }
#pragma unroll
            for (int i = KB; i < 9; ++i)
#pragma unroll
                for (int j = 0; j <= KB; ++j) x[i][j] = ::fma(-l[i], w[j], x[i][j]);
        }
    }
    __device__ inline void diag_linv_reg(int nd, const double* __restrict__ D, double* __restrict__ Li,
                                         double* __restrict__ bufA, double* __restrict__ bufB)
    {
        const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
        double x[9][9];
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j) x[i][j] = (tx + 16 * i == ty + 16 * j) ? 1.0 : 0.0;
        double nxt = (tid < nd) ? D[tid] : 0.0; // column k of D, loaded one step ahead
        linv_block_steps<0>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<1>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<2>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<3>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<4>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<5>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<6>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<7>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
        linv_block_steps<8>(nd, D, x, nxt, bufA, bufB, tx, ty, tid);
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j)
            {
                const int r = tx + 16 * i, c = ty + 16 * j;
                if (r < nd && c < nd) Li[r + static_cast<long long>(c) * nd] = x[i][j];
            }
    }

    // U_II^-1 of the factored block D (upper part, with the diagonal): X = I;
    // for k descending, X_k /= U_kk, rows r < k: X_r -= U_rk X_k.  Row k of X is
    // zero left of column k, so a step of block KB updates rows of blocks <= KB
    // and columns of blocks >= KB; the diagonal entry and the column of U are
    // loaded one step ahead.
    template <int KB>
    __device__ __forceinline__ void uinv_block_steps(int nd, const double* __restrict__ D, double (&x)[9][9],
                                                     double& nxt, double& ukk, double* __restrict__ bufA,
                                                     double* __restrict__ bufB, int tx, int ty, int tid)
    {
#pragma unroll 1
        for (int kt = 15; kt >= 0; --kt)
        {
            const int k = 16 * KB + kt;
            if (k >= nd) continue;
            double* rowk = bufA + (k & 1) * GT;
            double* colk = bufB + (k & 1) * GT;
            if (tid < k) colk[tid] = nxt; // column k of U above the diagonal
            const double rukk = 1.0 / ukk;
            if (k > 0)
            {
                if (tid < nd) nxt = D[tid + static_cast<long long>(k - 1) * nd];
                ukk = D[(k - 1) + static_cast<long long>(k - 1) * nd];
            }
            if (tx == kt)
#pragma unroll
                for (int j = KB; j < 9; ++j)
                {
                    x[KB][j] *= rukk;
                    rowk[ty + 16 * j] = x[KB][j];
                }
            __syncthreads();
            double w[9], u[9];
#pragma unroll
            for (int j = KB; j < 9; ++j) w[j] = rowk[ty + 16 * j];
#pragma unroll
            for (int i = 0; i <= KB; ++i)
            {
                const double v = colk[tx + 16 * i];
                u[i]           = (i < KB || tx < kt) ? v : 0.0;
            }
#pragma unroll
            for (int i = 0; i <= KB; ++i)
#pragma unroll
                for (int j = KB; j < 9; ++j) x[i][j] = ::fma(-u[i], w[j], x[i][j]);
        }
    }
    __device__ inline void diag_uinv_reg(int nd, const double* __restrict__ D, double* __restrict__ Ui,
                                         double* __restrict__ bufA, double* __restrict__ bufB)
    {
        const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
        double x[9][9];
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j) x[i][j] = (tx + 16 * i == ty + 16 * j) ? 1.0 : 0.0;
        double nxt = (tid < nd) ? D[tid + static_cast<long long>(nd - 1) * nd] : 0.0;
        double ukk = D[(nd - 1) + static_cast<long long>(nd - 1) * nd];
        uinv_block_steps<8>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<7>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<6>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<5>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<4>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<3>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<2>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<1>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
        uinv_block_steps<0>(nd, D, x, nxt, ukk, bufA, bufB, tx, ty, tid);
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j)
            {
                const int r = tx + 16 * i, c = ty + 16 * j;
                if (r < nd && c < nd) Ui[r + static_cast<long long>(c) * nd] = x[i][j];
            }
    }

    // The same for blocks larger than 144, in global memory: the LU, then one
    // inverse (which = 0: L_II^-1, 1: U_II^-1) with one column per thread.
    __device__ inline void diag_lu_global(int nd, double* D, int* info)
    {
        const int tid = threadIdx.x;
        for (int k = 0; k < nd; ++k)
        {
            const double piv = D[k + static_cast<long long>(k) * nd];
            if (::fabs(piv) <= 1e-12 && tid == 0) info[0] = 1; // PETSc's default zero-pivot tolerance
            for (int r = k + 1 + tid; r < nd; r += blockDim.x) D[r + static_cast<long long>(k) * nd] /= piv;
            __syncthreads();
            const int m = nd - k - 1;
            for (long long o = tid; o < static_cast<long long>(m) * m; o += blockDim.x)
            {
                const int r = k + 1 + static_cast<int>(o % m), c = k + 1 + static_cast<int>(o / m);
                D[r + static_cast<long long>(c) * nd] -= D[r + static_cast<long long>(k) * nd] * D[k + static_cast<long long>(c) * nd];
            }
            __syncthreads();
        }
    }
    __device__ inline void diag_inv_global(int nd, const double* D, double* X, int which)
    {
        for (int c = threadIdx.x; c < nd; c += blockDim.x)
        {
            const long long cc = static_cast<long long>(c) * nd;
            if (which == 0)
            {
                for (int r = 0; r < c; ++r) X[r + cc] = 0.0;
                X[c + cc] = 1.0;
                for (int r = c + 1; r < nd; ++r)
                {
                    double acc = 0.0;
#pragma unroll 8
                    for (int k = c; k < r; ++k) acc -= D[r + static_cast<long long>(k) * nd] * X[k + cc];
                    X[r + cc] = acc;
                }
            }
            else
            {
                for (int r = c + 1; r < nd; ++r) X[r + cc] = 0.0;
                X[c + cc] = 1.0 / D[c + cc];
                for (int r = c - 1; r >= 0; --r)
                {
                    double acc = 0.0;
#pragma unroll 8
                    for (int k = r + 1; k <= c; ++k) acc -= D[r + static_cast<long long>(k) * nd] * X[k + cc];
                    X[r + cc] = acc / D[r + static_cast<long long>(r) * nd];
                }
            }
        }
    }

    // ---- tiled factorisation: several workgroups per block row ----------------
    //
    // A level holds few block rows, and one workgroup per row leaves most CUs
    // idle while each row runs a chain of dependent block products.  The
    // tiled path runs a level as a sequence of launches whose workgroups are
    // 48 x 48 output tiles of those products: for each lower neighbour K in
    // order, T1 = A_IK U_KK^-1 (k_ilu_lstep), then A_IJ -= T1 U_KJ for every
    // update pair together with the store of T1 as L_IK (k_ilu_pstep); then
    // the diagonal block (k_ilu_diag: one workgroup per row); then
    // T2_J = L_II^-1 A_IJ for the upper blocks (k_ilu_ustep) and their store
    // (k_ilu_ucopy).  Scratch per row position b: T1 at T + b (1 + CMAXN)
    // blocks, T2_q after it.
    //
    // A tile is nine 16 x 16 sub-tiles on the FP64 matrix cores
    // (v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
    // B[l / 16][l % 16] of a 16x4 by 4x16 product and holds D[l / 16 + 4 v][l % 16]
    // in accumulator v); wave w takes sub-tiles w, w + 4, w + 8.  The cores
    // compute the transposed product B^T A^T, so that a lane's results run
    // down a column of the column-major block and the stores stay coalesced.
    // The k dimension is staged through LDS in 16-deep slabs, the next slab
    // loaded into registers while the cores work on the current one.
    constexpr int TT = 48, TKB = 16;
    typedef double dbl4 __attribute__((ext_vector_type(4)));
    __device__ inline void tile_gemm(int nd, int tr, int tc, const double* __restrict__ A,
                                     const double* __restrict__ B, double* __restrict__ C, bool sub,
                                     double* __restrict__ As, double* __restrict__ Bs)
    {
        const int tid = threadIdx.x, w = tid >> 6, li = tid & 15, lk = (tid >> 4) & 3;
        const int r0 = tr * TT, c0 = tc * TT;
        dbl4 acc[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) acc[m] = dbl4{0.0, 0.0, 0.0, 0.0};
        // TT x TKB values of A (rows r0.., k) and TKB x TT of B (k, columns c0..) per slab, FM each per thread
        constexpr int FM = TT * TKB / 256;
        double pa[FM], pb[FM];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int m = 0; m < FM; ++m)
            {
                const int o = tid + 256 * m;
                const int q = o / TT, s = o - q * TT;    // A: column k0 + q of the tile rows, s down the column
                const int r = r0 + s, ka = k0 + q;
                pa[m] = (r < nd && ka < nd) ? A[r + static_cast<long long>(ka) * nd] : 0.0;
                const int cq = o / TKB, cs = o - cq * TKB; // B: column c0 + cq, cs along k
                const int c = c0 + cq, kb = k0 + cs;
                pb[m] = (c < nd && kb < nd) ? B[kb + static_cast<long long>(c) * nd] : 0.0;
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < nd; k0 += TKB)
        {
#pragma unroll
            for (int m = 0; m < FM; ++m)
            {
                const int o = tid + 256 * m;
                const int q = o / TT, s = o - q * TT, cq = o / TKB, cs = o - cq * TKB;
                As[q * TT + s]   = pa[m]; // As[k][row]
                Bs[cs * TT + cq] = pb[m]; // Bs[k][column]
            }
            __syncthreads();
            if (k0 + TKB < nd) fetch(k0 + TKB);
#pragma unroll
            for (int ks = 0; ks < TKB; ks += 4)
            {
#pragma unroll
                for (int m = 0; m < 3; ++m)
                {
                    const int t = w + 4 * m;
                    if (t < 9) // uniform per wave
                    {
                        const int ti = t % 3, tj = t / 3;
                        const double bt = Bs[(ks + lk) * TT + 16 * tj + li]; // (B^T)[l % 16][l / 16]
                        const double at = As[(ks + lk) * TT + 16 * ti + li]; // (A^T)[l / 16][l % 16]
                        acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(bt, at, acc[m], 0, 0, 0);
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int m = 0; m < 3; ++m)
        {
            const int t = w + 4 * m;
            if (t >= 9) continue;
            const int ti = t % 3, tj = t / 3;
            const int r = r0 + 16 * ti + li;
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int c = c0 + 16 * tj + lk + 4 * v; // (B^T A^T)[l / 16 + 4 v][l % 16] = (A B)[r][c]
                if (r < nd && c < nd)
                {
                    const long long at = r + static_cast<long long>(c) * nd;
                    C[at] = sub ? C[at] - acc[m][v] : acc[m][v];
                }
            }
        }
    }

    __device__ inline void tile_copy(int nd, int tr, int tc, const double* __restrict__ src, double* __restrict__ dst)
    {
        for (int o = threadIdx.x; o < TT * TT; o += 256)
        {
            const int cc = o / TT, rr = o - cc * TT;
            const int r = tr * TT + rr, c = tc * TT + cc;
            if (r < nd && c < nd) dst[r + static_cast<long long>(c) * nd] = src[r + static_cast<long long>(c) * nd];
        }
    }

    // blockIdx.x: row position in the level; blockIdx.y: output tile
    static __global__ __launch_bounds__(256) void k_ilu_lstep(int S, int nd, int q, const int* __restrict__ lpos,
                                                              const int* __restrict__ lvl, int off,
                                                              const double* __restrict__ P,
                                                              const double* __restrict__ Dinv, double* __restrict__ T)
    {
        __shared__ double As[TT * TKB], Bs[TT * TKB];
        const long long pos = off + blockIdx.x;
        const int       sK  = lpos[pos * 2 * CMAXN + q];
        if (sK < 0) return;
        const long long K = lpos[pos * 2 * CMAXN + CMAXN + q], I = lvl[pos];
        const long long bsz = static_cast<long long>(nd) * nd;
        const int nt = (nd + TT - 1) / TT, tr = blockIdx.y % nt, tc = blockIdx.y / nt;
        tile_gemm(nd, tr, tc, P + (I * S + sK) * bsz, Dinv + K * 2 * bsz + bsz,
                  T + static_cast<long long>(blockIdx.x) * (1 + CMAXN) * bsz, false, As, Bs);
    }

    // blockIdx.y = task * nt^2 + tile; tasks 0..npair-1 the update pairs (npair:
    // the most any row of the level has for this neighbour), task npair the
    // store of T1 as L_IK
    static __global__ __launch_bounds__(256) void k_ilu_pstep(int S, int nd, int q, int npair, const int* __restrict__ lpos,
                                                              const int* __restrict__ upd,
                                                              const int* __restrict__ lvl, int off, double* P,
                                                              const double* __restrict__ T)
    {
        __shared__ double As[TT * TKB], Bs[TT * TKB];
        const long long pos = off + blockIdx.x;
        const int       sK  = lpos[pos * 2 * CMAXN + q];
        if (sK < 0) return;
        const long long K = lpos[pos * 2 * CMAXN + CMAXN + q], I = lvl[pos];
        const long long bsz = static_cast<long long>(nd) * nd;
        const int nt = (nd + TT - 1) / TT, nt2 = nt * nt;
        const int task = blockIdx.y / nt2, tile = blockIdx.y - task * nt2, tr = tile % nt, tc = tile / nt;
        const double* L = T + static_cast<long long>(blockIdx.x) * (1 + CMAXN) * bsz;
        if (task == npair)
        {
            tile_copy(nd, tr, tc, L, P + (I * S + sK) * bsz);
            return;
        }
        const long long at = ((I * CMAXN + q) * CMAXP + task) * 2;
        const int sj = upd[at];
        if (sj < 0) return;
        tile_gemm(nd, tr, tc, L, P + (K * S + upd[at + 1]) * bsz, P + (I * S + sj) * bsz, true, As, Bs);
    }

    // REG: blocks of at most 144 (register tile), else the global-memory path
    template <bool REG>
    __global__ __launch_bounds__(256) void k_ilu_diag(int S, int nd, const int* __restrict__ lvl, int off, double* P,
                                                      int* __restrict__ info)
    {
        __shared__ double bufA[2 * GT], bufB[2 * GT];
        const long long I = lvl[off + blockIdx.x];
        double* D = P + (I * S + S / 2) * static_cast<long long>(nd) * nd;
        if constexpr (REG)
            diag_lu_reg(nd, D, bufA, bufB, info);
        else
            diag_lu_global(nd, D, info);
    }

    // the two triangular inverses of the factored diagonal blocks, one
    // workgroup each (blockIdx.y = 0: L_II^-1, 1: U_II^-1)
    template <bool REG>
    __global__ __launch_bounds__(256) void k_ilu_inv(int S, int nd, const int* __restrict__ lvl, int off,
                                                     const double* __restrict__ P, double* __restrict__ Dinv)
    {
        __shared__ double bufA[2 * GT], bufB[2 * GT];
        const long long I   = lvl[off + blockIdx.x];
        const long long bsz = static_cast<long long>(nd) * nd;
        const double*   D   = P + (I * S + S / 2) * bsz;
        double*         X   = Dinv + I * 2 * bsz + (blockIdx.y ? bsz : 0);
        if constexpr (REG)
        {
            if (blockIdx.y == 0)
                diag_linv_reg(nd, D, X, bufA, bufB);
            else
                diag_uinv_reg(nd, D, X, bufA, bufB);
        }
        else
            diag_inv_global(nd, D, X, blockIdx.y);
    }

    // blockIdx.y = q * nt^2 + tile: T2_q = L_II^-1 A_IJ for the q-th upper block
    static __global__ __launch_bounds__(256) void k_ilu_ustep(int S, int nd, const int* __restrict__ upos,
                                                              const int* __restrict__ lvl, int off,
                                                              const double* __restrict__ P,
                                                              const double* __restrict__ Dinv, double* __restrict__ T)
    {
        __shared__ double As[TT * TKB], Bs[TT * TKB];
        const long long pos = off + blockIdx.x;
        const long long bsz = static_cast<long long>(nd) * nd;
        const int nt = (nd + TT - 1) / TT, nt2 = nt * nt;
        const int q = blockIdx.y / nt2, tile = blockIdx.y - q * nt2, tr = tile % nt, tc = tile / nt;
        const int sJ = upos[pos * 2 * CMAXN + q];
        if (sJ < 0) return;
        const long long I = lvl[pos];
        tile_gemm(nd, tr, tc, Dinv + I * 2 * bsz, P + (I * S + sJ) * bsz,
                  T + (static_cast<long long>(blockIdx.x) * (1 + CMAXN) + 1 + q) * bsz, false, As, Bs);
    }

    static __global__ __launch_bounds__(256) void k_ilu_ucopy(int S, int nd, const int* __restrict__ upos,
                                                              const int* __restrict__ lvl, int off, double* P,
                                                              const double* __restrict__ T)
    {
        const long long pos = off + blockIdx.x;
        const long long bsz = static_cast<long long>(nd) * nd;
        const int nt = (nd + TT - 1) / TT, nt2 = nt * nt;
        const int q = blockIdx.y / nt2, tile = blockIdx.y - q * nt2, tr = tile % nt, tc = tile / nt;
        const int sJ = upos[pos * 2 * CMAXN + q];
        if (sJ < 0) return;
        const long long I = lvl[pos];
        tile_copy(nd, tr, tc, T + (static_cast<long long>(blockIdx.x) * (1 + CMAXN) + 1 + q) * bsz,
                  P + (I * S + sJ) * bsz);
    }

    // After the last level: every off-diagonal block is replaced by its
    // product with the row's diagonal inverse, M_IK = L_II^-1 L_IK below and
    // N_IJ = U_II^-1 U_IJ above, so that the sweeps' rows are independent:
    // y_I = L_II^-1 x_I - sum_K M_IK y_K, z_I = U_II^-1 y_I - sum_J N_IJ z_J.
    // Blocks (row, off-diagonal slot) are taken in chunks starting at `first`
    // (blockIdx.x: block of the chunk, blockIdx.y: 48 x 48 tile); phase 0
    // copies the chunk's blocks to the scratch T, phase 1 writes the products.
    static __global__ __launch_bounds__(256) void k_ilu_scale(int S, int nd, long long nblk, long long first, int phase,
                                                              const int* __restrict__ nbr, double* P,
                                                              const double* __restrict__ Dinv, double* T)
    {
        __shared__ double As[TT * TKB], Bs[TT * TKB];
        const long long g = first + blockIdx.x;
        if (g >= nblk) return; // uniform per workgroup
        const long long e = g / (S - 1);
        int             s = static_cast<int>(g - e * (S - 1));
        if (s >= S / 2) ++s; // skip the centre slot
        const int m = nbr[e * S + s];
        if (m < 0) return;
        const long long bsz = static_cast<long long>(nd) * nd;
        const int nt = (nd + TT - 1) / TT, tr = blockIdx.y % nt, tc = blockIdx.y / nt;
        double* B  = P + (e * S + s) * bsz;
        double* Tw = T + static_cast<long long>(blockIdx.x) * bsz;
        if (phase == 0)
            tile_copy(nd, tr, tc, B, Tw);
        else
            tile_gemm(nd, tr, tc, Dinv + e * 2 * bsz + (m < e ? 0 : bsz), Tw, B, false, As, Bs);
    }

    constexpr int CMAXND = 336; // largest element block (2D p = 8: 324 DOFs)

    // One level of the forward (out_I = L_II^-1 in_I - sum_K M_IK out_K, K
    // below) or backward (out_I = U_II^-1 in_I - sum_J N_IJ out_J, J above)
    // sweep on the scaled factors.  Rows are independent, so a row is split
    // over workgroups: blockIdx.y picks 16 rows, thread (row, g) sums the
    // columns c = g (mod 16) of every block and 16 partial sums meet in LDS.
    // A level holds few rows; spreading each over nd/16 workgroups puts the
    // whole chip on it.  ND = block size as a template parameter (0: run time)
    // so that a thread's loads are all in flight together.
    constexpr int SWR = 16;       // rows per workgroup
    constexpr int SWG = 256 / SWR; // column groups: thread (row, g) sums columns g, g + SWG, ...
    template <int ND, class T>
    __device__ inline double sweep_dot(int nd, int r, int g, const T* __restrict__ B,
                                       const double* __restrict__ v)
    {
        double acc = 0.0;
        if constexpr (ND > 0)
        {
#pragma unroll
            for (int c = g; c < ND; c += SWG) acc = ::fma(static_cast<double>(B[r + c * ND]), v[c], acc);
        }
        else
        {
            for (int c = g; c < nd; c += SWG)
                acc = ::fma(static_cast<double>(B[r + static_cast<long long>(c) * nd]), v[c], acc);
        }
        return acc;
    }

    // T: the stored factors' type (double; float for the Newton preconditioner's
    // single-precision copy, accumulated in double)
    template <bool FWD, int ND, class T = double>
    __global__ __launch_bounds__(256) void k_ilu_sweep(int S, int nd, const int* __restrict__ pos_tab,
                                                       const int* __restrict__ lvl, int off,
                                                       const T* __restrict__ P,
                                                       const T* __restrict__ Dinv,
                                                       const double* __restrict__ in, double* out)
    {
        __shared__ double red[SWG][SWR + 1];
        if constexpr (ND > 0) nd = ND;
        const long long pos = off + blockIdx.x;
        const long long I   = lvl[pos];
        const long long bsz = static_cast<long long>(nd) * nd;
        const int       rl  = threadIdx.x % SWR, g = threadIdx.x / SWR;
        const int       r   = blockIdx.y * SWR + rl;
        int sl[CMAXN], el[CMAXN];
#pragma unroll
        for (int q = 0; q < CMAXN; ++q)
        {
            sl[q] = pos_tab[pos * 2 * CMAXN + q];
            el[q] = pos_tab[pos * 2 * CMAXN + CMAXN + q];
        }
        double acc = 0.0;
        if constexpr (ND > 0)
        {
            // The row's blocks (the diagonal inverse, then the neighbours'
            // scaled blocks, packed in pos_tab) in batches of SWB: every load of
            // a batch is issued before its first FMA, so a row pays one memory
            // latency per batch rather than one per block.  Addresses are
            // clamped into the block (rows and columns past ND are discarded or
            // multiplied by zero), so no load is masked.  Per thread the sums
            // run in the same order as sweep_dot's.
            constexpr int NC = (ND + SWG - 1) / SWG, SWB = 5;
            int nb = 1;
#pragma unroll
            for (int q = 0; q < CMAXN; ++q) nb += sl[q] >= 0 ? 1 : 0;
            const int rr = r < ND ? r : ND - 1;
#pragma unroll
            for (int b0 = 0; b0 <= CMAXN; b0 += SWB)
            {
                if (b0 >= nb) break; // uniform per workgroup
                double bv[SWB][NC], vv[SWB][NC];
#pragma unroll
                for (int t = 0; t < SWB; ++t)
                {
                    const int qq = b0 + t;
                    if (qq > CMAXN) break;
                    if (qq < nb)
                    {
                        const T*      B = qq == 0 ? Dinv + I * 2 * bsz + (FWD ? 0 : bsz) : P + (I * S + sl[qq - 1]) * bsz;
                        const double* V = qq == 0 ? in + I * nd : out + static_cast<long long>(el[qq - 1]) * nd;
#pragma unroll
                        for (int i = 0; i < NC; ++i)
                        {
                            const int c = g + SWG * i, cc = c < ND ? c : ND - 1;
                            bv[t][i] = static_cast<double>(B[rr + cc * ND]);
                            const double x = V[cc];
                            vv[t][i] = c < ND ? x : 0.0;
                        }
                    }
                    else
                    {
#pragma unroll
                        for (int i = 0; i < NC; ++i) bv[t][i] = vv[t][i] = 0.0;
                    }
                }
#pragma unroll
                for (int t = 0; t < SWB; ++t)
                {
                    const int qq = b0 + t;
                    if (qq > CMAXN) break;
                    double part = 0.0;
#pragma unroll
                    for (int i = 0; i < NC; ++i) part = ::fma(bv[t][i], vv[t][i], part);
                    if (qq == 0)
                        acc = part;
                    else if (qq < nb)
                        acc -= part;
                }
            }
        }
        else if (r < nd)
        {
            acc = sweep_dot<ND>(nd, r, g, Dinv + I * 2 * bsz + (FWD ? 0 : bsz), in + I * nd);
#pragma unroll
            for (int q = 0; q < CMAXN; ++q)
                if (sl[q] >= 0)
                    acc -= sweep_dot<ND>(nd, r, g, P + (I * S + sl[q]) * bsz, out + static_cast<long long>(el[q]) * nd);
        }
        red[g][rl] = acc;
        __syncthreads();
        if (g == 0 && r < nd)
        {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < SWG; ++k) sum += red[k][rl];
            out[I * nd + r] = sum;
        }
    }

    // One sweep direction of k_ilu_sweep in a single launch over every block
    // row, ordered by data dependence instead of by level launches.  A
    // workgroup draws a ticket from *ctr when it starts; ticket t is chunk
    // t % RCH of the block row at level-order position t / RCH (from the end
    // for the backward sweep).  It loads its first SWB blocks and its own
    // input, then waits until every chunk of each neighbour row it reads has
    // been stored (done[el] == RCH), sums exactly as k_ilu_sweep does (same
    // per-thread order, so the same results) and counts its own chunk into
    // done[I].  A neighbour row sits at an earlier level, so its chunks hold
    // earlier tickets: their workgroups were already running when this one
    // drew, and every wait ends without a grid barrier or co-residency
    // assumption.  As a guard, a wait longer than DF_SPIN polls, or one that
    // finds *err set, sets *err and gives up (the host reports it); every
    // workgroup therefore exits.  ctr, done and err start at zero (reset by
    // k_ilu_df_reset before each apply).
    //
    // Visibility without cache maintenance: out rows are padded to whole
    // 128-byte lines (stride NDP) and stored at device scope (written
    // through to the point of coherence; the wave waits for the stores before
    // it counts the chunk), and a row is read only after all its chunks were
    // counted, so no cache line of a row is ever read before the row is
    // final and plain loads see the stored values.  Release / acquire fences
    // per workgroup (a whole-L2 write-back and invalidation each) made this
    // kernel 4x slower than the level launches.
    constexpr int DF_SPIN = 1 << 22;
    __host__ __device__ constexpr int df_stride(int nd) { return (nd + 15) / 16 * 16; }
    __global__ void k_ilu_df_reset(int n, int* __restrict__ w)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) w[i] = 0;
    }
    template <int ND>
    __global__ void k_ilu_df_unpad(long long E, const double* __restrict__ zp, double* __restrict__ z)
    {
        const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i >= E * ND) return;
        const long long e = i / ND;
        z[i]              = zp[e * df_stride(ND) + (i - e * ND)];
    }
    template <bool FWD, int ND, class T>
    __global__ __launch_bounds__(256) void k_ilu_sweep_df(int S, int E, const int* __restrict__ pos_tab,
                                                          const int* __restrict__ lvl, const T* __restrict__ P,
                                                          const T* __restrict__ Dinv, const double* __restrict__ in,
                                                          double* out, int* ctr, int* done, int* err)
    {
        static_assert(ND > 0 && ND <= 144, "compile-time block size up to 144");
        constexpr int NC = (ND + SWG - 1) / SWG, RCH = (ND + SWR - 1) / SWR, SWB = 5;
        constexpr int NDP = df_stride(ND), IS = FWD ? ND : NDP; // out / in row strides
        constexpr long long bsz = static_cast<long long>(ND) * ND;
        __shared__ double red[SWG][SWR + 1];
        __shared__ int    tk;
        if (threadIdx.x == 0) tk = atomicAdd(ctr, 1);
        __syncthreads();
        const int tkt = tk;
        if (tkt >= E * RCH) return; // uniform; not reached with E * RCH workgroups
        const int       p   = tkt / RCH;
        const long long pos = FWD ? p : E - 1 - p;
        const long long I   = lvl[pos];
        const int       rl  = threadIdx.x % SWR, g = threadIdx.x / SWR;
        const int       r   = (tkt - p * RCH) * SWR + rl;
        int sl[CMAXN], el[CMAXN];
#pragma unroll
        for (int q = 0; q < CMAXN; ++q)
        {
            sl[q] = pos_tab[pos * 2 * CMAXN + q];
            el[q] = pos_tab[pos * 2 * CMAXN + CMAXN + q];
        }
        int nb = 1;
#pragma unroll
        for (int q = 0; q < CMAXN; ++q) nb += sl[q] >= 0 ? 1 : 0;
        const int rr = r < ND ? r : ND - 1;
        auto blk = [&](int qq) -> const T*
        { return qq == 0 ? Dinv + I * 2 * bsz + (FWD ? 0 : bsz) : P + (I * S + sl[qq - 1]) * bsz; };
        // the first batch of blocks and the row's own input: nothing to wait for
        double bv[SWB][NC], vv[SWB][NC];
#pragma unroll
        for (int t = 0; t < SWB; ++t)
        {
            if (t < nb)
            {
                const T* B = blk(t);
#pragma unroll
                for (int i = 0; i < NC; ++i)
                {
                    const int c = g + SWG * i, cc = c < ND ? c : ND - 1;
                    bv[t][i]    = static_cast<double>(B[rr + cc * ND]);
                }
            }
            else
#pragma unroll
                for (int i = 0; i < NC; ++i) bv[t][i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < NC; ++i)
        {
            const int    c = g + SWG * i, cc = c < ND ? c : ND - 1;
            const double x = in[I * IS + cc];
            vv[0][i]       = c < ND ? x : 0.0;
        }
        // the diagonal block's part first (it is the first term of the sum)
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NC; ++i) acc = ::fma(bv[0][i], vv[0][i], acc);
        if (threadIdx.x == 0)
        {
            for (int q = 0; q + 1 < nb; ++q)
            {
                const int* f = done + el[q];
                int        n = 0;
                while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < RCH)
                {
                    if (++n > DF_SPIN || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    {
                        atomicOr(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int b0 = 0; b0 <= CMAXN; b0 += SWB)
        {
            if (b0 >= nb) break; // uniform per workgroup
#pragma unroll
            for (int t = 0; t < SWB; ++t)
            {
                const int qq = b0 + t;
                if (qq > CMAXN) break;
                if (b0 > 0)
                {
                    if (qq < nb)
                    {
                        const T* B = blk(qq);
#pragma unroll
                        for (int i = 0; i < NC; ++i)
                        {
                            const int c = g + SWG * i, cc = c < ND ? c : ND - 1;
                            bv[t][i]    = static_cast<double>(B[rr + cc * ND]);
                        }
                    }
                    else
#pragma unroll
                        for (int i = 0; i < NC; ++i) bv[t][i] = 0.0;
                }
                if (qq == 0) continue; // the row's own input, loaded above
                if (qq < nb)
                {
                    const double* V = out + static_cast<long long>(el[qq - 1]) * NDP;
#pragma unroll
                    for (int i = 0; i < NC; ++i)
                    {
                        const int    c = g + SWG * i, cc = c < ND ? c : ND - 1;
                        const double x = V[cc];
                        vv[t][i]       = c < ND ? x : 0.0;
                    }
                }
                else
#pragma unroll
                    for (int i = 0; i < NC; ++i) vv[t][i] = 0.0;
            }
#pragma unroll
            for (int t = 0; t < SWB; ++t)
            {
                const int qq = b0 + t;
                if (qq > CMAXN) break;
                if (qq == 0) continue;
                double part = 0.0;
#pragma unroll
                for (int i = 0; i < NC; ++i) part = ::fma(bv[t][i], vv[t][i], part);
                if (qq < nb) acc -= part;
            }
        }
        red[g][rl] = acc;
        __syncthreads();
        if (g == 0 && r < ND)
        {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < SWG; ++k) sum += red[k][rl];
            __hip_atomic_store(out + I * NDP + r, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the stores above come from lanes 0..15 of wave 0: that wave waits
        // until they have completed, then counts the chunk as stored
        if (threadIdx.x < 64)
        {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (threadIdx.x == 0) __hip_atomic_fetch_add(done + I, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // Transposed apply, z = (L U)^-T x, on the same stored factors (the
    // adjoint's preconditioner).  With t_J = U_JJ^T y_J and s_K = L_KK^T z_K:
    //   forward levels:  t_I = x_I - sum_{J below} N_JI^T t_J,  y_I = U_II^-T t_I
    //   backward levels: s_I = y_I - sum_{K above} M_KI^T s_K,  z_I = L_II^-T s_I
    // where N_JI / M_KI are the neighbour row's scaled blocks at the mirrored
    // slot S - 1 - s (the box stencil's offsets are symmetric).  One element per
    // workgroup of 16 waves: wave w takes rows w + 16 i, lanes take columns
    // (the transposed blocks are read along their contiguous columns), and a
    // row's 64 lane sums are reduced by cross-lane shuffles.
    template <int ND>
    __device__ inline void sweepT_block(int nd, int lane, int w, const double* __restrict__ B,
                                        const double* __restrict__ v, double (&acc)[21])
    {
        constexpr int RW = ND ? (ND + 15) / 16 : 21;
        if constexpr (ND > 0)
        {
#pragma unroll
            for (int c0 = 0; c0 < ND; c0 += 64)
            {
                const int  c  = c0 + lane;
                const bool ok = c < ND;
                const double vc = ok ? v[c] : 0.0;
#pragma unroll
                for (int i = 0; i < RW; ++i)
                {
                    const int r = w + 16 * i;
                    if (ok && r < ND) acc[i] = ::fma(B[c + r * ND], vc, acc[i]);
                }
            }
        }
        else
        {
            for (int c = lane; c < nd; c += 64)
            {
                const double vc = v[c];
#pragma unroll
                for (int i = 0; i < RW; ++i)
                {
                    const int r = w + 16 * i;
                    if (r < nd) acc[i] = ::fma(B[c + static_cast<long long>(r) * nd], vc, acc[i]);
                }
            }
        }
    }

    __device__ inline double wave_sum(double v)
    {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    }

    template <bool FWD, int ND>
    __global__ __launch_bounds__(1024) void k_ilu_sweepT(int S, int nd, const int* __restrict__ pos_tab,
                                                         const int* __restrict__ lvl, int off,
                                                         const double* __restrict__ P,
                                                         const double* __restrict__ Dinv,
                                                         const double* __restrict__ in, double* aux, double* out)
    {
        __shared__ double tl[CMAXND];
        if constexpr (ND > 0) nd = ND;
        constexpr int RW = ND ? (ND + 15) / 16 : 21;
        const long long pos  = off + blockIdx.x;
        const long long I    = lvl[pos];
        const long long bsz  = static_cast<long long>(nd) * nd;
        const int       lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        int sl[CMAXN], el[CMAXN];
#pragma unroll
        for (int q = 0; q < CMAXN; ++q)
        {
            sl[q] = pos_tab[pos * 2 * CMAXN + q];
            el[q] = pos_tab[pos * 2 * CMAXN + CMAXN + q];
        }
        double acc[21];
#pragma unroll
        for (int i = 0; i < 21; ++i) acc[i] = 0.0;
#pragma unroll
        for (int q = 0; q < CMAXN; ++q)
            if (sl[q] >= 0)
                sweepT_block<ND>(nd, lane, w, P + (static_cast<long long>(el[q]) * S + (S - 1 - sl[q])) * bsz,
                                 aux + static_cast<long long>(el[q]) * nd, acc);
#pragma unroll
        for (int i = 0; i < RW; ++i)
        {
            const double a = wave_sum(acc[i]);
            const int    r = w + 16 * i;
            if (lane == 0 && r < nd)
            {
                const double tv = in[I * nd + r] - a;
                tl[r]           = tv;
                aux[I * nd + r] = tv;
            }
            acc[i] = 0.0;
        }
        __syncthreads();
        // out_I = (U_II^-1)^T t_I (forward) or (L_II^-1)^T s_I (backward)
        sweepT_block<ND>(nd, lane, w, Dinv + I * 2 * bsz + (FWD ? bsz : 0), tl, acc);
#pragma unroll
        for (int i = 0; i < RW; ++i)
        {
            const double a = wave_sum(acc[i]);
            const int    r = w + 16 * i;
            if (lane == 0 && r < nd) out[I * nd + r] = a;
        }
    }

    // The level recurrences of the transposed sweeps need only the neighbours'
    // t (forward) or s (backward), not their diagonal products, so those
    // products leave the level chain: each level runs k_ilu_sweepT_lvl,
    //   aux_I = in_I - sum_q B_q^T aux_{m_q}   (B_q: row m_q's block at the mirrored slot),
    // with a block row split over workgroups (16 entries each, 4 per wave,
    // lanes along the blocks' contiguous columns) and the loads of SWB blocks
    // issued before their FMAs; then one launch over all rows forms
    // out_I = (U_II^-1)^T aux_I or (L_II^-1)^T aux_I (k_ilu_diagT).  Each
    // entry's lane partial sums and their shuffle reduction follow
    // k_ilu_sweepT's order, so the results are the same.
    template <int ND>
    __global__ __launch_bounds__(256) void k_ilu_sweepT_lvl(int S, const int* __restrict__ pos_tab,
                                                            const int* __restrict__ lvl, int off,
                                                            const double* __restrict__ P,
                                                            const double* __restrict__ in, double* aux)
    {
        constexpr int NCH = (ND + 63) / 64, SWB = ND <= 144 ? 4 : (ND <= 256 ? 2 : 1);
        constexpr long long bsz = static_cast<long long>(ND) * ND;
        const long long pos = off + blockIdx.x;
        const long long I   = lvl[pos];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, rb = blockIdx.y * 16 + w * 4;
        int sl[CMAXN], el[CMAXN];
#pragma unroll
        for (int q = 0; q < CMAXN; ++q)
        {
            sl[q] = pos_tab[pos * 2 * CMAXN + q];
            el[q] = pos_tab[pos * 2 * CMAXN + CMAXN + q];
        }
        int nb = 0; // slots are packed: the row's neighbours first, then -1
#pragma unroll
        for (int q = 0; q < CMAXN; ++q) nb += sl[q] >= 0 ? 1 : 0;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int b0 = 0; b0 < CMAXN; b0 += SWB)
        {
            if (b0 >= nb) break; // uniform per workgroup
            double bv[SWB][4][NCH], vv[SWB][NCH];
#pragma unroll
            for (int t = 0; t < SWB; ++t)
            {
                const int q = b0 + t;
                if (q < nb)
                {
                    const double* B = P + (static_cast<long long>(el[q]) * S + (S - 1 - sl[q])) * bsz;
                    const double* V = aux + static_cast<long long>(el[q]) * ND;
#pragma unroll
                    for (int ch = 0; ch < NCH; ++ch)
                    {
                        const int c = ch * 64 + lane, cc = c < ND ? c : ND - 1;
                        const double x = V[cc];
                        vv[t][ch] = c < ND ? x : 0.0;
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                        {
                            const int r = rb + j, rr = r < ND ? r : ND - 1;
                            bv[t][j][ch] = B[cc + rr * ND];
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < SWB; ++t)
                if (b0 + t < nb)
#pragma unroll
                    for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[j] = ::fma(bv[t][j][ch], vv[t][ch], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const double a = wave_sum(acc[j]);
            const int    r = rb + j;
            if (lane == 0 && r < ND) aux[I * ND + r] = in[I * ND + r] - a;
        }
    }

    // blockIdx.x: block row, blockIdx.y: 16 entries
    template <bool FWD, int ND>
    __global__ __launch_bounds__(256) void k_ilu_diagT(const double* __restrict__ Dinv, const double* __restrict__ aux,
                                                       double* out)
    {
        constexpr int NCH = (ND + 63) / 64;
        constexpr long long bsz = static_cast<long long>(ND) * ND;
        const long long I = blockIdx.x;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, rb = blockIdx.y * 16 + w * 4;
        const double* B = Dinv + I * 2 * bsz + (FWD ? bsz : 0);
        double bv[4][NCH], vv[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
        {
            const int c = ch * 64 + lane, cc = c < ND ? c : ND - 1;
            const double x = aux[I * ND + cc];
            vv[ch] = c < ND ? x : 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int r = rb + j, rr = r < ND ? r : ND - 1;
                bv[j][ch] = B[cc + rr * ND];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            double acc = 0.0;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc = ::fma(bv[j][ch], vv[ch], acc);
            const double a = wave_sum(acc);
            const int    r = rb + j;
            if (lane == 0 && r < ND) out[I * ND + r] = a;
        }
    }
} // namespace frsdg

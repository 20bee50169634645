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
// whose lower neighbours are all factored form a level; levels run as
// separate launches (i + 2j for a non-periodic 2D mesh: nx + 2 ny - 2 levels).
// The diagonal blocks' triangular inverses are kept so that both sweeps of
// the apply are dense block products.
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
        std::vector<int> cls_count; // elements per colour class (host)
        int* nbr  = nullptr; // [E][S] neighbour element per slot, -1 none
        int* cls  = nullptr; // [E] colour class
        int* low  = nullptr; // [E][CMAXN] lower slots (neighbour index < e) in increasing neighbour order
        int* upd  = nullptr; // [E][CMAXN][2 CMAXP] (slot of J in row I, slot of J in row K) per lower slot
        int* upp  = nullptr; // [E][CMAXN] upper slots
        int* lvl  = nullptr; // [E] elements grouped by level
        int* info = nullptr; // zero-pivot flag
        double* P    = nullptr; // E*S*nd*nd, factored in place by the preconditioner
        double* Dinv = nullptr; // E*2*nd*nd: L_II^-1, U_II^-1
        double* y    = nullptr; // forward-sweep vector
        double* v    = nullptr; // perturbed state / unit colour vector
        double* f0   = nullptr; // F at the base state
        double* f1   = nullptr; // F at the perturbed state
    };

    // Neighbour tables, colour classes and the ILU(0) level schedule.
    // Returns FRSDG_ERR_UNSUPPORTED where PETSc's DMDA colouring refuses the
    // mesh (a periodic direction whose element count is not a multiple of 3).
    inline int color_plan_tables(int dims, int nx, int ny, int per_x, int per_y, int nd, ColorPlan& cp,
                                 std::vector<int>& h_nbr, std::vector<int>& h_cls, std::vector<int>& h_low,
                                 std::vector<int>& h_upd, std::vector<int>& h_upp, std::vector<int>& h_lvl)
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
                                                                const int* __restrict__ cls, int cc, int kdof,
                                                                const double* __restrict__ f1,
                                                                const double* __restrict__ f0,
                                                                const double* __restrict__ x, double dxinv_wp,
                                                                double eps, double umin, int mode,
                                                                double* __restrict__ P)
    {
        const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (idx >= E * nd) return;
        const long long R = idx / nd;
        const int r = static_cast<int>(idx - R * nd);
        for (int s = 0; s < S; ++s)
        {
            const int m = nbr[R * S + s];
            if (m < 0 || cls[m] != cc) continue;
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
            break; // one neighbour of a colour class per block row
        }
    }

    // ---- block ILU(0) -----------------------------------------------------------

    // One block row per workgroup, rows of one level per launch.
    static __global__ __launch_bounds__(256) void k_ilu_row(int S, int nd, const int* __restrict__ nbr,
                                                            const int* __restrict__ low, const int* __restrict__ upd,
                                                            const int* __restrict__ upp, const int* __restrict__ lvl,
                                                            int off, double* P, double* __restrict__ Dinv,
                                                            int* __restrict__ info)
    {
        const long long I   = lvl[off + blockIdx.x];
        const long long bsz = static_cast<long long>(nd) * nd;
        const int       C   = S / 2; // centre slot
        const int       tid = threadIdx.x;
        auto blk = [&](long long e, int s) { return P + (e * S + s) * bsz; };
        for (int q = 0; q < CMAXN; ++q)
        {
            const int sK = low[I * CMAXN + q];
            if (sK < 0) break;
            const long long K = nbr[I * S + sK];
            double*       A = blk(I, sK);
            const double* U = blk(K, C);
            // L_IK = A_IK U_KK^-1, each row by forward substitution over the columns
            for (int r = tid; r < nd; r += blockDim.x)
                for (int c = 0; c < nd; ++c)
                {
                    double acc = A[r + static_cast<long long>(c) * nd];
                    for (int k = 0; k < c; ++k) acc -= A[r + static_cast<long long>(k) * nd] * U[k + static_cast<long long>(c) * nd];
                    A[r + static_cast<long long>(c) * nd] = acc / U[c + static_cast<long long>(c) * nd];
                }
            __syncthreads();
            for (int p = 0; p < CMAXP; ++p)
            {
                const long long at = ((I * CMAXN + q) * CMAXP + p) * 2;
                const int sj = upd[at];
                if (sj < 0) break;
                double*       B  = blk(I, sj);
                const double* UK = blk(K, upd[at + 1]);
                for (long long o = tid; o < bsz; o += blockDim.x)
                {
                    const int r = static_cast<int>(o % nd), c = static_cast<int>(o / nd);
                    double acc = B[o];
                    for (int k = 0; k < nd; ++k) acc -= A[r + static_cast<long long>(k) * nd] * UK[k + static_cast<long long>(c) * nd];
                    B[o] = acc;
                }
                __syncthreads();
            }
        }
        // A_II = L_II U_II, unpivoted, right-looking
        double* D = blk(I, C);
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
        // U_IJ = L_II^-1 A_IJ, each column by forward substitution
        for (int q = 0; q < CMAXN; ++q)
        {
            const int sJ = upp[I * CMAXN + q];
            if (sJ < 0) break;
            double* B = blk(I, sJ);
            for (int c = tid; c < nd; c += blockDim.x)
                for (int r = 1; r < nd; ++r)
                {
                    double acc = B[r + static_cast<long long>(c) * nd];
                    for (int k = 0; k < r; ++k) acc -= D[r + static_cast<long long>(k) * nd] * B[k + static_cast<long long>(c) * nd];
                    B[r + static_cast<long long>(c) * nd] = acc;
                }
        }
        // triangular inverses of the diagonal block, one column per thread
        double* Li = Dinv + I * 2 * bsz;
        double* Ui = Li + bsz;
        for (int c = tid; c < nd; c += blockDim.x)
        {
            const long long cc = static_cast<long long>(c) * nd;
            for (int r = 0; r < c; ++r) Li[r + cc] = 0.0;
            Li[c + cc] = 1.0;
            for (int r = c + 1; r < nd; ++r)
            {
                double acc = 0.0;
                for (int k = c; k < r; ++k) acc -= D[r + static_cast<long long>(k) * nd] * Li[k + cc];
                Li[r + cc] = acc;
            }
            for (int r = c + 1; r < nd; ++r) Ui[r + cc] = 0.0;
            Ui[c + cc] = 1.0 / D[c + cc];
            for (int r = c - 1; r >= 0; --r)
            {
                double acc = 0.0;
                for (int k = r + 1; k <= c; ++k) acc -= D[r + static_cast<long long>(k) * nd] * Ui[k + cc];
                Ui[r + cc] = acc / D[r + static_cast<long long>(r) * nd];
            }
        }
    }

    constexpr int CMAXND = 336; // largest element block (2D p = 8: 324 DOFs)

    // forward sweep of one level: y_I = L_II^-1 (x_I - sum_K L_IK y_K)
    static __global__ __launch_bounds__(256) void k_ilu_fwd(int S, int nd, const int* __restrict__ nbr,
                                                            const int* __restrict__ low, const int* __restrict__ lvl,
                                                            int off, const double* __restrict__ P,
                                                            const double* __restrict__ Dinv,
                                                            const double* __restrict__ x, double* y)
    {
        __shared__ double t[CMAXND];
        const long long I   = lvl[off + blockIdx.x];
        const long long bsz = static_cast<long long>(nd) * nd;
        for (int r = threadIdx.x; r < nd; r += blockDim.x)
        {
            double acc = x[I * nd + r];
            for (int q = 0; q < CMAXN; ++q)
            {
                const int sK = low[I * CMAXN + q];
                if (sK < 0) break;
                const long long K  = nbr[I * S + sK];
                const double*   L  = P + (I * S + sK) * bsz;
                const double*   yk = y + K * nd;
                for (int c = 0; c < nd; ++c) acc -= L[r + static_cast<long long>(c) * nd] * yk[c];
            }
            t[r] = acc;
        }
        __syncthreads();
        const double* Li = Dinv + I * 2 * bsz;
        for (int r = threadIdx.x; r < nd; r += blockDim.x)
        {
            double acc = 0.0;
            for (int c = 0; c <= r; ++c) acc += Li[r + static_cast<long long>(c) * nd] * t[c];
            y[I * nd + r] = acc;
        }
    }

    // backward sweep of one level: z_I = U_II^-1 (y_I - sum_J U_IJ z_J)
    static __global__ __launch_bounds__(256) void k_ilu_bwd(int S, int nd, const int* __restrict__ nbr,
                                                            const int* __restrict__ upp, const int* __restrict__ lvl,
                                                            int off, const double* __restrict__ P,
                                                            const double* __restrict__ Dinv,
                                                            const double* __restrict__ y, double* z)
    {
        __shared__ double t[CMAXND];
        const long long I   = lvl[off + blockIdx.x];
        const long long bsz = static_cast<long long>(nd) * nd;
        for (int r = threadIdx.x; r < nd; r += blockDim.x)
        {
            double acc = y[I * nd + r];
            for (int q = 0; q < CMAXN; ++q)
            {
                const int sJ = upp[I * CMAXN + q];
                if (sJ < 0) break;
                const long long J  = nbr[I * S + sJ];
                const double*   U  = P + (I * S + sJ) * bsz;
                const double*   zj = z + J * nd;
                for (int c = 0; c < nd; ++c) acc -= U[r + static_cast<long long>(c) * nd] * zj[c];
            }
            t[r] = acc;
        }
        __syncthreads();
        const double* Ui = Dinv + I * 2 * bsz + bsz;
        for (int r = threadIdx.x; r < nd; r += blockDim.x)
        {
            double acc = 0.0;
            for (int c = r; c < nd; ++c) acc += Ui[r + static_cast<long long>(c) * nd] * t[c];
            z[I * nd + r] = acc;
        }
    }
} // namespace frsdg

// frsdg_newton.hip — Jacobian-free Newton-Krylov for the implicit (IFunction)
// path: F(u) = (u - u^n)/dt + R(u), J = (1/dt) I + dR/du applied matrix-free
// by the forward-tangent or complex-step kernel (the reference's MatShell,
// petsc_solver_impl.hpp:168-359, driven by PETSc SNES newtonls + KSP GMRES).
//
// Device pieces: deterministic two-pass reductions (block partials, then a
// fixed-order sum), classical Gram-Schmidt against the Krylov basis in one
// pass (as PETSc's default KSPGMRESClassicalGramSchmidtOrthogonalization),
// and an element block-Jacobi preconditioner whose exact diagonal blocks are
// probed by 2 * (4 n^2) tangent products on a checkerboard colouring (face
// neighbours never share a colour, so each probe isolates one column of every
// diagonal block of its colour) and inverted in place by Gauss-Jordan.
// The small Hessenberg least-squares problem runs on the host (Givens).
#include "frsdg_color.hpp"
#include "frsdg_ctx.hpp"

#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

using namespace frsdg;

struct SolverWs
{
    long long N   = 0;
    int       m   = 0;       // restart length the basis was sized for
    double*   V   = nullptr; // (m+1) x N Krylov basis
    double*   Z   = nullptr; // m x N preconditioned basis of flexible GMRES (allocated on first use)
    int       mZ  = 0;
    double*   w   = nullptr; // work vector
    double*   z   = nullptr; // preconditioned vector / update
    double*   un  = nullptr; // u^n
    double*   ud  = nullptr; // udot
    double*   F   = nullptr; // nonlinear residual
    double*   du  = nullptr; // Newton update
    double*   ut  = nullptr; // line-search trial state
    double*   part = nullptr; // reduction partials
    double*   red  = nullptr; // reduced scalars (device)
    int       nb   = 0;       // reduction blocks
    // block-Jacobi
    double* B    = nullptr; // E blocks of (nb_e x nb_e), column-major, inverted in place
    float*  Bf   = nullptr; // FP32 copy of the inverses used by the preconditioner apply
    double* pv   = nullptr; // probe vector
    double* pjv  = nullptr; // probe product
    int*    info = nullptr; // per-block singularity flag
    // coloured preconditioning matrix and its ILU(0)
    frsdg::ColorPlan* cp = nullptr;
    // lagged preconditioner of the Newton solves: which one the buffers hold
    // (-1 none / stale) and how many Newton iterations have used it
    int pc_built = -1;
    int pc_age   = 0;
    // timing events of the per-part breakdown (frsdg_newton_stats t_*)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace frsdg
{
    void free_solver(frsdg_ctx* c)
    {
        if (!c || !c->ws) return;
        SolverWs* s = c->ws;
        for (double* p : {s->V, s->Z, s->w, s->z, s->un, s->ud, s->F, s->du, s->ut, s->part, s->red, s->B, s->pv, s->pjv})
            if (p) (void) hipFree(p);
        if (s->info) (void) hipFree(s->info);
        if (s->Bf) (void) hipFree(s->Bf);
        for (hipEvent_t& e : s->ev)
            if (e) (void) hipEventDestroy(e);
        if (frsdg::ColorPlan* cp = s->cp)
        {
            for (int* p : {cp->nbr, cp->cls, cp->cslot, cp->low, cp->upd, cp->upp, cp->lvl, cp->lpos, cp->upos, cp->info, cp->df})
                if (p) (void) hipFree(p);
            for (double* p : {cp->P, cp->Dinv, cp->y, cp->v, cp->f0, cp->f1, cp->T, cp->t, cp->ydf, cp->zdf})
                if (p) (void) hipFree(p);
            for (float* p : {cp->P32, cp->Dinv32})
                if (p) (void) hipFree(p);
            delete cp;
        }
        delete s;
        c->ws = nullptr;
    }
} // namespace frsdg

namespace
{
    constexpr int RB = 256; // reduction block size

    // partial dot products of k vectors V_j (stride N) with w: part[j][block]
    __global__ __launch_bounds__(RB) void k_mdot(long long N, const double* __restrict__ V, int k,
                                                 const double* __restrict__ w, int nb, double* __restrict__ part)
    {
        __shared__ double s[RB];
        const int j = blockIdx.y;
        const double* v = V + j * N;
        double acc = 0.0;
        for (long long i = blockIdx.x * static_cast<long long>(RB) + threadIdx.x; i < N;
             i += static_cast<long long>(nb) * RB)
            acc += v[i] * w[i];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int st = RB / 2; st > 0; st >>= 1)
        {
            if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[static_cast<long long>(j) * nb + blockIdx.x] = s[0];
        (void) k;
    }
    // out[j] = sum_b part[j][b] (fixed order)
    __global__ __launch_bounds__(RB) void k_reduce(int nb, const double* __restrict__ part, double* __restrict__ out)
    {
        __shared__ double s[RB];
        const int j = blockIdx.x;
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += RB) acc += part[static_cast<long long>(j) * nb + b];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int st = RB / 2; st > 0; st >>= 1)
        {
            if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = s[0];
    }
    // w -= sum_j h[j] V_j ; part[b] = partial ||w||^2
    __global__ __launch_bounds__(RB) void k_maxpy_norm(long long N, const double* __restrict__ V, int k,
                                                       const double* __restrict__ h, double* __restrict__ w,
                                                       int nb, double* __restrict__ part)
    {
        __shared__ double s[RB];
        double acc = 0.0;
        for (long long i = blockIdx.x * static_cast<long long>(RB) + threadIdx.x; i < N;
             i += static_cast<long long>(nb) * RB)
        {
            double x = w[i];
            for (int j = 0; j < k; ++j) x -= h[j] * V[j * N + i];
            w[i] = x;
            acc += x * x;
        }
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int st = RB / 2; st > 0; st >>= 1)
        {
            if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x] = s[0];
    }
    // y = a*x + b*y
    __global__ void k_axpby(long long N, double a, const double* __restrict__ x, double b, double* __restrict__ y)
    {
        const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (i < N) y[i] = a * x[i] + b * y[i];
    }
    // z = a*x + b*y
    __global__ void k_waxpby(long long N, double a, const double* __restrict__ x, double b,
                             const double* __restrict__ y, double* __restrict__ z)
    {
        const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (i < N) z[i] = a * x[i] + b * y[i];
    }
    // y = sum_j c[j] V_j (c on device)
    __global__ void k_mcomb(long long N, const double* __restrict__ V, int k, const double* __restrict__ c,
                            double* __restrict__ y)
    {
        const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (i >= N) return;
        double x = 0.0;
        for (int j = 0; j < k; ++j) x += c[j] * V[j * N + i];
        y[i] = x;
    }

    // ---------------- element block-Jacobi ----------------
    // probe vector: 1 at local DOF m of every element of colour col (checkerboard), else 0
    __global__ void k_probe(long long E, int nx, int nd_e, int m, int col, double* __restrict__ v)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= E * nd_e) return;
        const long long e = t / nd_e;
        const int q = static_cast<int>(t - e * nd_e);
        const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        v[t] = (q == m && ((i + j) & 1) == col) ? 1.0 : 0.0;
    }
    // column m of the diagonal blocks of colour col: B_e[:, m] = (J v)_e
    __global__ void k_probe_store(long long E, int nx, int nd_e, int m, int col, const double* __restrict__ jv,
                                  double* __restrict__ B)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= E * nd_e) return;
        const long long e = t / nd_e;
        const int r = static_cast<int>(t - e * nd_e);
        const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        if (((i + j) & 1) != col) return;
        B[e * nd_e * nd_e + static_cast<long long>(m) * nd_e + r] = jv[t];
    }
    // in-place Gauss-Jordan inverse of each block (column-major, partial pivoting
    // on the row index), one workgroup per block; the block stays in global
    // memory (4 n^2 up to 324: 0.8 MB) and is cached.
    __global__ __launch_bounds__(256) void k_block_inverse(int nd_e, double* __restrict__ Ball, int* __restrict__ info)
    {
        const long long e = blockIdx.x;
        double* A = Ball + e * nd_e * nd_e; // A[c*nd + r]
        __shared__ int piv_s;
        __shared__ double pval[256];
        __shared__ int pidx[256];
        __shared__ int perm[336];
        for (int r = threadIdx.x; r < nd_e; r += blockDim.x) perm[r] = r;
        if (threadIdx.x == 0) info[e] = 0;
        __syncthreads();
        for (int c = 0; c < nd_e; ++c)
        {
            // pivot search over rows c..nd-1 of column c
            double best = -1.0;
            int bi = c;
            for (int r = c + threadIdx.x; r < nd_e; r += blockDim.x)
            {
                const double a = ::fabs(A[static_cast<long long>(c) * nd_e + r]);
                if (a > best) { best = a; bi = r; }
            }
            pval[threadIdx.x] = best;
            pidx[threadIdx.x] = bi;
            __syncthreads();
            for (int st = blockDim.x / 2; st > 0; st >>= 1)
            {
                if (threadIdx.x < st)
                {
                    const double o = pval[threadIdx.x + st];
                    const int oi = pidx[threadIdx.x + st];
                    if (o > pval[threadIdx.x] || (o == pval[threadIdx.x] && oi < pidx[threadIdx.x]))
                    {
                        pval[threadIdx.x] = o;
                        pidx[threadIdx.x] = oi;
                    }
                }
                __syncthreads();
            }
            if (threadIdx.x == 0)
            {
                piv_s = pidx[0];
                if (!(pval[0] > 0.0)) info[e] = 1;
            }
            __syncthreads();
            const int p = piv_s;
            // swap rows c and p in all columns
            if (p != c)
            {
                for (int cc = threadIdx.x; cc < nd_e; cc += blockDim.x)
                {
                    double* col = A + static_cast<long long>(cc) * nd_e;
                    const double tmp = col[c];
                    col[c] = col[p];
                    col[p] = tmp;
                }
                if (threadIdx.x == 0)
                {
                    const int t = perm[c];
                    perm[c] = perm[p];
                    perm[p] = t;
                }
            }
            __syncthreads();
            const double d = A[static_cast<long long>(c) * nd_e + c];
            const double inv = (d != 0.0) ? 1.0 / d : 0.0;
            // scale row c
            for (int cc = threadIdx.x; cc < nd_e; cc += blockDim.x)
            {
                if (cc == c) continue;
                A[static_cast<long long>(cc) * nd_e + c] *= inv;
            }
            __syncthreads();
            // eliminate column c from other rows: A[r][cc] -= A[r][c] * A[c][cc]; A[r][c] = -A[r][c] * inv
            for (long long t = threadIdx.x; t < static_cast<long long>(nd_e) * nd_e; t += blockDim.x)
            {
                const int cc = static_cast<int>(t / nd_e), r = static_cast<int>(t - static_cast<long long>(cc) * nd_e);
                if (r == c || cc == c) continue;
                A[static_cast<long long>(cc) * nd_e + r] -= A[static_cast<long long>(c) * nd_e + r] * A[static_cast<long long>(cc) * nd_e + c];
            }
            __syncthreads();
            for (int r = threadIdx.x; r < nd_e; r += blockDim.x)
            {
                if (r == c) continue;
                A[static_cast<long long>(c) * nd_e + r] = -A[static_cast<long long>(c) * nd_e + r] * inv;
            }
            if (threadIdx.x == 0) A[static_cast<long long>(c) * nd_e + c] = inv;
            __syncthreads();
        }
        // undo the row permutation as a column permutation of the inverse:
        // inv(P A) = inv(A) P^T  ->  inv(A) = inv(PA) P
        // columns of inv(A): col perm[k] of inv(A) = col k of inv(PA)
        __shared__ double tmpc[336];
        for (int r = 0; r < nd_e; ++r)
        {
            for (int k = threadIdx.x; k < nd_e; k += blockDim.x) tmpc[perm[k]] = A[static_cast<long long>(k) * nd_e + r];
            __syncthreads();
            for (int k = threadIdx.x; k < nd_e; k += blockDim.x) A[static_cast<long long>(k) * nd_e + r] = tmpc[k];
            __syncthreads();
        }
    }
    __global__ void k_to_float(long long n, const double* __restrict__ a, float* __restrict__ b)
    {
        const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (i < n) b[i] = static_cast<float>(a[i]);
    }
    // z_e = Binv_e x_e (column-major block: coalesced over rows), or Binv_e^T x_e.
    // The inverses are stored in FP32 (a preconditioner; halves its HBM traffic),
    // the products accumulate in FP64.
    __global__ void k_block_apply(long long E, int nd_e, const float* __restrict__ Binv, const double* __restrict__ x,
                                  double* __restrict__ z, int transpose)
    {
        const long long e = blockIdx.x;
        const float*  A  = Binv + e * nd_e * nd_e;
        const double* xe = x + e * nd_e;
        for (int r = threadIdx.x; r < nd_e; r += blockDim.x)
        {
            double s = 0.0;
            if (transpose)
                for (int c = 0; c < nd_e; ++c) s += A[static_cast<long long>(r) * nd_e + c] * xe[c];
            else
                for (int c = 0; c < nd_e; ++c) s += A[static_cast<long long>(c) * nd_e + r] * xe[c];
            z[e * nd_e + r] = s;
        }
        (void) E;
    }

    // z_e = B_e^-1 x_e on the elements of colour col (checkerboard (i + j) & 1);
    // the other colour is zeroed when zero_other is set, left as is otherwise
    __global__ void k_block_apply_color(long long E, int nx, int nd_e, const float* __restrict__ Binv,
                                        const double* __restrict__ x, double* __restrict__ z, int transpose, int col,
                                        int zero_other)
    {
        const long long e = blockIdx.x;
        const int       i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        if (((i + j) & 1) != col)
        {
            if (zero_other)
                for (int r = threadIdx.x; r < nd_e; r += blockDim.x) z[e * nd_e + r] = 0.0;
            return;
        }
        const float*  A  = Binv + e * nd_e * nd_e;
        const double* xe = x + e * nd_e;
        for (int r = threadIdx.x; r < nd_e; r += blockDim.x)
        {
            double s = 0.0;
            if (transpose)
                for (int c = 0; c < nd_e; ++c) s += A[static_cast<long long>(r) * nd_e + c] * xe[c];
            else
                for (int c = 0; c < nd_e; ++c) s += A[static_cast<long long>(c) * nd_e + r] * xe[c];
            z[e * nd_e + r] = s;
        }
        (void) E;
    }

    // Tables and buffers of the coloured matrix and its ILU(0), once per context.
    int color_plan_ensure(frsdg_ctx* c, SolverWs* s)
    {
        if (s->cp) return FRSDG_OK;
        const int dims = c->d.dims;
        const int nd   = static_cast<int>(c->nd / c->E);
        if (nd > frsdg::CMAXND) return FRSDG_ERR_UNSUPPORTED;
        auto* cp = new frsdg::ColorPlan{};
        std::vector<int> nbr, cls, low, upd, upp, lvl, lpos, upos;
        int rc = frsdg::color_plan_tables(dims, c->d.nx, dims == 2 ? c->d.ny : 1, dims == 2 ? c->p2.per_x : c->p1.periodic,
                                          dims == 2 ? c->p2.per_y : 0, nd, *cp, nbr, cls, low, upd, upp, lvl, lpos, upos);
        if (rc)
        {
            delete cp;
            return rc;
        }
        s->cp = cp; // owned by the workspace from here on, also if an allocation below fails
        auto up = [](int*& d, const std::vector<int>& h) -> int
        {
            FR_TRY(hipMalloc(&d, sizeof(int) * std::max<size_t>(1, h.size())));
            FR_TRY(hipMemcpy(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
            return FRSDG_OK;
        };
        // slot of each block row's neighbour in each colour class
        std::vector<int> cslot(static_cast<size_t>(cp->ncls) * cp->E, -1);
        for (long long e = 0; e < cp->E; ++e)
            for (int sl = 0; sl < cp->S; ++sl)
            {
                const int m = nbr[e * cp->S + sl];
                if (m >= 0) cslot[static_cast<size_t>(cls[m]) * cp->E + e] = sl;
            }
        if ((rc = up(cp->nbr, nbr)) || (rc = up(cp->cls, cls)) || (rc = up(cp->cslot, cslot)) || (rc = up(cp->low, low)) ||
            (rc = up(cp->upd, upd)) || (rc = up(cp->upp, upp)) || (rc = up(cp->lvl, lvl)) || (rc = up(cp->lpos, lpos)) ||
            (rc = up(cp->upos, upos)))
            return rc;
        FR_TRY(hipMalloc(&cp->info, sizeof(int)));
        FR_TRY(hipMalloc(&cp->df, sizeof(int) * (3 + 2 * cp->E)));
        FR_TRY(hipMemset(cp->df, 0, sizeof(int) * (3 + 2 * cp->E)));
        for (double** p : {&cp->ydf, &cp->zdf}) FR_TRY(hipMalloc(p, sizeof(double) * cp->E * frsdg::df_stride(nd)));
        const long long bsz = static_cast<long long>(nd) * nd;
        int width = 1;
        for (size_t l = 0; l + 1 < cp->lvl_ptr.size(); ++l) width = std::max(width, cp->lvl_ptr[l + 1] - cp->lvl_ptr[l]);
        // scratch blocks: 1 + CMAXN per row of the widest level, at least 256 for the scaling pass
        cp->tcap = std::max<long long>(static_cast<long long>(width) * (1 + frsdg::CMAXN), 256);
        FR_TRY(hipMalloc(&cp->T, sizeof(double) * cp->tcap * bsz));
        for (double** p : {&cp->y, &cp->v, &cp->f0, &cp->f1}) FR_TRY(hipMalloc(p, sizeof(double) * s->N));
        return FRSDG_OK;
    }

    // Block ILU(0) of P in place (frsdg_color.hpp), level by level.
    int ilu_factor(SolverWs* s, hipStream_t st, double* P, double* Dinv)
    {
        const frsdg::ColorPlan& cp = *s->cp;
        FR_TRY(hipMemsetAsync(cp.info, 0, sizeof(int), st));
        const int nlev = static_cast<int>(cp.lvl_ptr.size()) - 1;
        const unsigned nt2 = static_cast<unsigned>(((cp.nd + frsdg::TT - 1) / frsdg::TT) * ((cp.nd + frsdg::TT - 1) / frsdg::TT));
        for (int l = 0; l < nlev; ++l)
        {
            const unsigned rows = static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l]);
            const int off = cp.lvl_ptr[l];
            for (int q = 0; q < cp.lvl_nlow[l]; ++q)
            {
                hipLaunchKernelGGL(frsdg::k_ilu_lstep, dim3(rows, nt2), dim3(256), 0, st, cp.S, cp.nd, q, cp.lpos, cp.lvl,
                                   off, P, Dinv, cp.T);
                const int np = cp.lvl_npair[static_cast<size_t>(l) * frsdg::CMAXN + q];
                hipLaunchKernelGGL(frsdg::k_ilu_pstep, dim3(rows, static_cast<unsigned>(np + 1) * nt2), dim3(256), 0, st,
                                   cp.S, cp.nd, q, np, cp.lpos, cp.upd, cp.lvl, off, P, cp.T);
            }
            if (cp.nd <= frsdg::GT)
            {
                hipLaunchKernelGGL(frsdg::k_ilu_diag<true>, dim3(rows), dim3(256), 0, st, cp.S, cp.nd, cp.lvl, off, P,
                                   cp.info);
                hipLaunchKernelGGL(frsdg::k_ilu_inv<true>, dim3(rows, 2), dim3(256), 0, st, cp.S, cp.nd, cp.lvl, off, P,
                                   Dinv);
            }
            else
            {
                hipLaunchKernelGGL(frsdg::k_ilu_diag<false>, dim3(rows), dim3(256), 0, st, cp.S, cp.nd, cp.lvl, off, P,
                                   cp.info);
                hipLaunchKernelGGL(frsdg::k_ilu_inv<false>, dim3(rows, 2), dim3(256), 0, st, cp.S, cp.nd, cp.lvl, off, P,
                                   Dinv);
            }
            if (cp.lvl_nupp[l] > 0)
            {
                const dim3 g(rows, static_cast<unsigned>(cp.lvl_nupp[l]) * nt2);
                hipLaunchKernelGGL(frsdg::k_ilu_ustep, g, dim3(256), 0, st, cp.S, cp.nd, cp.upos, cp.lvl, off, P, Dinv, cp.T);
                hipLaunchKernelGGL(frsdg::k_ilu_ucopy, g, dim3(256), 0, st, cp.S, cp.nd, cp.upos, cp.lvl, off, P, cp.T);
            }
        }
        // scale the off-diagonal blocks (rows independent in the sweeps), in
        // chunks of as many workgroups as the scratch holds blocks
        const long long nblk = cp.E * (cp.S - 1);
        for (long long first = 0; first < nblk; first += cp.tcap)
        {
            const dim3 g(static_cast<unsigned>(std::min<long long>(cp.tcap, nblk - first)), nt2);
            hipLaunchKernelGGL(frsdg::k_ilu_scale, g, dim3(256), 0, st, cp.S, cp.nd, nblk, first, 0, cp.nbr, P, Dinv, cp.T);
            hipLaunchKernelGGL(frsdg::k_ilu_scale, g, dim3(256), 0, st, cp.S, cp.nd, nblk, first, 1, cp.nbr, P, Dinv, cp.T);
        }
        FR_TRY(hipGetLastError());
        int info = 0;
        FR_TRY(hipMemcpyAsync(&info, cp.info, sizeof(int), hipMemcpyDeviceToHost, st));
        FR_TRY(hipStreamSynchronize(st));
        return info ? FRSDG_ERR_NUMERIC : FRSDG_OK;
    }

    // z = (L U)^-1 x: forward sweep over the levels, then backward (z may alias x);
    // compile-time block sizes run each direction as one dependence-ordered launch
    template <int ND, class T = double>
    void ilu_sweeps(const frsdg::ColorPlan& cp, hipStream_t st, const T* P, const T* Dinv, const double* x,
                    double* z)
    {
        const int nlev = static_cast<int>(cp.lvl_ptr.size()) - 1;
        const unsigned rch = static_cast<unsigned>((cp.nd + frsdg::SWR - 1) / frsdg::SWR);
        if constexpr (ND > 0 && ND <= 144 && std::is_same_v<T, double>)
        {
            // one dependence-ordered launch per direction (k_ilu_sweep_df; the
            // single-precision factors keep the level launches, 6.5 vs 7.9 ms,
            // and so do blocks above 144, whose registers allow one wave per
            // SIMD; at 144 two, and a cap at three spills and runs 2.4x slower)
            const int      E  = static_cast<int>(cp.E);
            const int      nw = 3 + 2 * E;
            const unsigned wg = static_cast<unsigned>(E) * rch;
            hipLaunchKernelGGL(frsdg::k_ilu_df_reset, dim3(static_cast<unsigned>((nw + 255) / 256)), dim3(256), 0, st,
                               nw, cp.df);
            hipLaunchKernelGGL((frsdg::k_ilu_sweep_df<true, ND, T>), dim3(wg), dim3(256), 0, st, cp.S, E, cp.lpos,
                               cp.lvl, P, Dinv, x, cp.ydf, cp.df, cp.df + 3, cp.df + 2);
            hipLaunchKernelGGL((frsdg::k_ilu_sweep_df<false, ND, T>), dim3(wg), dim3(256), 0, st, cp.S, E, cp.upos,
                               cp.lvl, P, Dinv, cp.ydf, cp.zdf, cp.df + 1, cp.df + 3 + E, cp.df + 2);
            hipLaunchKernelGGL((frsdg::k_ilu_df_unpad<ND>), dim3(static_cast<unsigned>((cp.E * ND + 255) / 256)), dim3(256),
                               0, st, cp.E, cp.zdf, z);
            return;
        }
        for (int l = 0; l < nlev; ++l)
            hipLaunchKernelGGL((frsdg::k_ilu_sweep<true, ND, T>),
                               dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l]), rch), dim3(256), 0, st, cp.S,
                               cp.nd, cp.lpos, cp.lvl, cp.lvl_ptr[l], P, Dinv, x, cp.y);
        for (int l = nlev - 1; l >= 0; --l)
            hipLaunchKernelGGL((frsdg::k_ilu_sweep<false, ND, T>),
                               dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l]), rch), dim3(256), 0, st, cp.S,
                               cp.nd, cp.upos, cp.lvl, cp.lvl_ptr[l], P, Dinv, cp.y, z);
    }
    // z = (L U)^-T x (the adjoint's preconditioner; z may alias x).
    template <int ND>
    void ilu_sweeps_T(const frsdg::ColorPlan& cp, hipStream_t st, const double* P, const double* Dinv, const double* x,
                      double* z)
    {
        const int nlev = static_cast<int>(cp.lvl_ptr.size()) - 1;
        if constexpr (ND > 0)
        {
            // level chains of t (then s) in cp.t; the diagonal products for all
            // rows after each chain (y into cp.y, then z: x is no longer read)
            const unsigned rch = static_cast<unsigned>((ND + 15) / 16);
            const dim3     all(static_cast<unsigned>(cp.E), rch);
            for (int l = 0; l < nlev; ++l)
                hipLaunchKernelGGL((frsdg::k_ilu_sweepT_lvl<ND>), dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l]), rch),
                                   dim3(256), 0, st, cp.S, cp.lpos, cp.lvl, cp.lvl_ptr[l], P, x, cp.t);
            hipLaunchKernelGGL((frsdg::k_ilu_diagT<true, ND>), all, dim3(256), 0, st, Dinv, cp.t, cp.y);
            for (int l = nlev - 1; l >= 0; --l)
                hipLaunchKernelGGL((frsdg::k_ilu_sweepT_lvl<ND>), dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l]), rch),
                                   dim3(256), 0, st, cp.S, cp.upos, cp.lvl, cp.lvl_ptr[l], P, cp.y, cp.t);
            hipLaunchKernelGGL((frsdg::k_ilu_diagT<false, ND>), all, dim3(256), 0, st, Dinv, cp.t, z);
        }
        else
        {
            for (int l = 0; l < nlev; ++l)
                hipLaunchKernelGGL((frsdg::k_ilu_sweepT<true, ND>), dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l])),
                                   dim3(1024), 0, st, cp.S, cp.nd, cp.lpos, cp.lvl, cp.lvl_ptr[l], P, Dinv, x, cp.t, cp.y);
            for (int l = nlev - 1; l >= 0; --l)
                hipLaunchKernelGGL((frsdg::k_ilu_sweepT<false, ND>), dim3(static_cast<unsigned>(cp.lvl_ptr[l + 1] - cp.lvl_ptr[l])),
                                   dim3(1024), 0, st, cp.S, cp.nd, cp.upos, cp.lvl, cp.lvl_ptr[l], P, Dinv, cp.y, cp.t, z);
        }
    }
    int ilu_apply_T(SolverWs* s, hipStream_t st, const double* P, const double* Dinv, const double* x, double* z)
    {
        frsdg::ColorPlan& cp = *s->cp;
        if (!cp.t) FR_TRY(hipMalloc(&cp.t, sizeof(double) * s->N));
        switch (cp.nd)
        {
            case 16: ilu_sweeps_T<16>(cp, st, P, Dinv, x, z); break;
            case 36: ilu_sweeps_T<36>(cp, st, P, Dinv, x, z); break;
            case 64: ilu_sweeps_T<64>(cp, st, P, Dinv, x, z); break;
            case 100: ilu_sweeps_T<100>(cp, st, P, Dinv, x, z); break;
            case 144: ilu_sweeps_T<144>(cp, st, P, Dinv, x, z); break;
            case 196: ilu_sweeps_T<196>(cp, st, P, Dinv, x, z); break;
            case 256: ilu_sweeps_T<256>(cp, st, P, Dinv, x, z); break;
            case 324: ilu_sweeps_T<324>(cp, st, P, Dinv, x, z); break;
            default: ilu_sweeps_T<0>(cp, st, P, Dinv, x, z); break;
        }
        return hip_rc(hipGetLastError());
    }

    template <class T>
    int ilu_apply(SolverWs* s, hipStream_t st, const T* P, const T* Dinv, const double* x, double* z)
    {
        const frsdg::ColorPlan& cp = *s->cp;
        switch (cp.nd)
        {
            // 2D element blocks 4 (p+1)^2, 1D 3 (p+1)
            case 16: ilu_sweeps<16, T>(cp, st, P, Dinv, x, z); break;
            case 36: ilu_sweeps<36, T>(cp, st, P, Dinv, x, z); break;
            case 64: ilu_sweeps<64, T>(cp, st, P, Dinv, x, z); break;
            case 100: ilu_sweeps<100, T>(cp, st, P, Dinv, x, z); break;
            case 144: ilu_sweeps<144, T>(cp, st, P, Dinv, x, z); break;
            case 196: ilu_sweeps<196, T>(cp, st, P, Dinv, x, z); break;
            case 256: ilu_sweeps<256, T>(cp, st, P, Dinv, x, z); break;
            case 324: ilu_sweeps<324, T>(cp, st, P, Dinv, x, z); break;
            default: ilu_sweeps<0, T>(cp, st, P, Dinv, x, z); break;
        }
        return hip_rc(hipGetLastError());
    }
    // FRSDG_ERR_HIP if a dependence-ordered sweep gave up a wait since the
    // last apply's reset (a broken wait protocol, never expected); synchronises
    int ilu_df_status(const frsdg::ColorPlan& cp, hipStream_t st)
    {
        int e = 0;
        FR_TRY(hipMemcpyAsync(&e, cp.df + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        FR_TRY(hipStreamSynchronize(st));
        return e ? FRSDG_ERR_HIP : FRSDG_OK;
    }

    struct Ctx
    {
        frsdg_ctx* c;
        SolverWs*  s;
        hipStream_t st;
        const frsdg_newton_opts* o;
        double t, sigma;
        const double* u; // linearisation point
        frsdg_newton_stats* stats;
        int transpose = 0; // solve with J^T (adjoint kernel) instead of J
        std::vector<double>* hist = nullptr; // GMRES residual norms: ||b||, then the estimate after each iteration

        unsigned grid(long long n) const { return static_cast<unsigned>((n + 255) / 256); }
        // seconds between two recorded events of s->ev (both complete)
        double between(int a, int b) const
        {
            float ms = 0.0f;
            return hipEventElapsedTime(&ms, s->ev[a], s->ev[b]) == hipSuccess ? 1e-3 * ms : 0.0;
        }
        void mark(int k) const { (void) hipEventRecord(s->ev[k], st); }

        int jv(const double* v, double* out)
        {
            stats->jv_evals++;
            if (transpose) return frsdg_vjp(c, t, u, v, sigma, out, st);
            if (o->jac_type == 1) return frsdg_jvp_cstep(c, t, u, v, sigma, 1e-7, out, st);
            return frsdg_jvp(c, t, u, v, sigma, out, st);
        }
        int prec(const double* x, double* z)
        {
            if (o->precond >= 3)
            {
                if (transpose) return ilu_apply_T(s, st, s->cp->P, s->cp->Dinv, x, z);
                if (o->pc_fp32 && s->cp->P32)
                    return ilu_apply(s, st, static_cast<const float*>(s->cp->P32),
                                     static_cast<const float*>(s->cp->Dinv32), x, z);
                return ilu_apply(s, st, static_cast<const double*>(s->cp->P), static_cast<const double*>(s->cp->Dinv),
                                 x, z);
            }
            if (o->precond == 2)
            {
                // red-black element block Gauss-Seidel, one forward sweep, matrix-free:
                // z_R = D_R^-1 x_R;  z_B = D_B^-1 (x - A z)_B  (A = sigma I + J, or its
                // transpose for the adjoint; A z restricted to black needs only the
                // red entries of z, which is all z holds at that point)
                const int nd_e = static_cast<int>(c->nd / c->E);
                const dim3 bs(nd_e <= 64 ? 64 : (nd_e <= 128 ? 128 : 256));
                hipLaunchKernelGGL(k_block_apply_color, dim3(static_cast<unsigned>(c->E)), bs, 0, st, c->E, c->d.nx,
                                   nd_e, s->Bf, x, z, transpose, 0, 1);
                int rc = jv(z, s->pjv);
                if (rc) return rc;
                hipLaunchKernelGGL(k_waxpby, dim3(grid(s->N)), dim3(256), 0, st, s->N, 1.0, x, -1.0, s->pjv, s->pv);
                hipLaunchKernelGGL(k_block_apply_color, dim3(static_cast<unsigned>(c->E)), bs, 0, st, c->E, c->d.nx,
                                   nd_e, s->Bf, s->pv, z, transpose, 1, 0);
                return hip_rc(hipGetLastError());
            }
            if (o->precond == 1)
            {
                const int nd_e = static_cast<int>(c->nd / c->E);
                hipLaunchKernelGGL(k_block_apply, dim3(static_cast<unsigned>(c->E)), dim3(nd_e <= 64 ? 64 : (nd_e <= 128 ? 128 : 256)), 0, st, c->E, nd_e, s->Bf, x, z, transpose);
                return hip_rc(hipGetLastError());
            }
            if (z != x) FR_TRY(hipMemcpyAsync(z, x, sizeof(double) * s->N, hipMemcpyDeviceToDevice, st));
            return FRSDG_OK;
        }
        // dot products of V_0..V_{k-1} with w -> s->red[0..k) (device)
        int mdot(const double* Vb, int k, const double* w)
        {
            hipLaunchKernelGGL(k_mdot, dim3(s->nb, k), dim3(RB), 0, st, s->N, Vb, k, w, s->nb, s->part);
            hipLaunchKernelGGL(k_reduce, dim3(k), dim3(RB), 0, st, s->nb, s->part, s->red);
            return hip_rc(hipGetLastError());
        }
        double norm(const double* w)
        {
            mdot(w, 1, w);
            double h = 0.0;
            (void) hipMemcpyAsync(&h, s->red, sizeof(double), hipMemcpyDeviceToHost, st);
            (void) hipStreamSynchronize(st);
            return std::sqrt(h);
        }
        // The coloured matrix dF/du of F at u (frsdg_color.hpp): method 0 coloured
        // finite differences with PETSc's default "wp" step, 1 the "ds" step,
        // 2 exact columns from tangent products of the unit colour vectors.
        template <class EvalF>
        int color_assemble(int method, EvalF&& evalF, double* P)
        {
            frsdg::ColorPlan& cp = *s->cp;
            const long long N = s->N;
            const double eps = 1.4901161193847656e-08; // sqrt(DBL_EPSILON): MatFDColoring error_rel
            const double umin = 100.0 * eps;             // MatFDColoring umin
            FR_TRY(hipMemsetAsync(P, 0, sizeof(double) * cp.E * cp.S * cp.nd * cp.nd, st));
            double h_wp = 0.0, dxinv = 0.0;
            if (method == 2)
                FR_TRY(hipMemsetAsync(cp.v, 0, sizeof(double) * N, st));
            else
            {
                int rc = evalF(u, cp.f0);
                if (rc) return rc;
                FR_TRY(hipMemcpyAsync(cp.v, u, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
                if (method == 0)
                {
                    const double unorm = norm(u);
                    dxinv = 1.0 / (std::sqrt(1.0 + unorm) * eps);
                    h_wp  = 1.0 / dxinv;
                }
            }
            const unsigned ge = static_cast<unsigned>((cp.E + 255) / 256);
            const int keep_t = transpose;
            for (int cc = 0; cc < cp.ncls; ++cc)
            {
                if (!cp.cls_count[cc]) continue;
                for (int k = 0; k < cp.nd; ++k)
                {
                    hipLaunchKernelGGL(frsdg::k_color_perturb, dim3(ge), dim3(256), 0, st, cp.E, cp.nd, cp.cls, cc, k, u,
                                       cp.v, h_wp, eps, umin, method, 0);
                    int rc;
                    if (method == 2)
                    {
                        transpose = 0;
                        rc        = jv(cp.v, cp.f1);
                        transpose = keep_t;
                    }
                    else
                        rc = evalF(cp.v, cp.f1);
                    if (rc) return rc;
                    // the store also restores the colour's entries of the perturbed copy
                    hipLaunchKernelGGL(frsdg::k_color_store, dim3(grid(N)), dim3(256), 0, st, cp.E, cp.S, cp.nd, cp.nbr,
                                       cp.cslot, cc, k, cp.f1, cp.f0, u, dxinv, eps, umin, method, P, cp.v);
                }
            }
            return hip_rc(hipGetLastError());
        }
        template <class EvalF>
        int build_ilu(int method, EvalF&& evalF)
        {
            int rc = color_plan_ensure(c, s);
            if (rc) return rc;
            frsdg::ColorPlan* cp = s->cp;
            if (!cp->P) // the preconditioner's own matrix, allocated on first use
            {
                const long long bsz = static_cast<long long>(cp->nd) * cp->nd;
                FR_TRY(hipMalloc(&cp->P, sizeof(double) * cp->E * cp->S * bsz));
                FR_TRY(hipMalloc(&cp->Dinv, sizeof(double) * cp->E * 2 * bsz));
            }
            mark(0);
            if ((rc = color_assemble(method, evalF, s->cp->P))) return rc;
            mark(1);
            if ((rc = ilu_factor(s, st, s->cp->P, s->cp->Dinv))) return rc;
            if (o->pc_fp32 && !transpose)
            {
                // single-precision copies of the factors for the applications
                const long long nP = cp->E * cp->S * static_cast<long long>(cp->nd) * cp->nd;
                const long long nD = cp->E * 2 * static_cast<long long>(cp->nd) * cp->nd;
                if (!cp->P32) FR_TRY(hipMalloc(&cp->P32, sizeof(float) * nP));
                if (!cp->Dinv32) FR_TRY(hipMalloc(&cp->Dinv32, sizeof(float) * nD));
                hipLaunchKernelGGL(k_to_float, dim3(grid(nP)), dim3(256), 0, st, nP, cp->P, cp->P32);
                hipLaunchKernelGGL(k_to_float, dim3(grid(nD)), dim3(256), 0, st, nD, cp->Dinv, cp->Dinv32);
                FR_TRY(hipGetLastError());
            }
            mark(2);
            FR_TRY(hipEventSynchronize(s->ev[2]));
            stats->t_pc_assemble += between(0, 1);
            stats->t_pc_factor += between(1, 2);
            return FRSDG_OK;
        }
        int build_block_jacobi()
        {
            // a 2-colour checkerboard separates face neighbours unless a
            // periodic direction has an odd element count
            if ((c->p2.per_x && (c->d.nx & 1)) || (c->p2.per_y && (c->d.ny & 1))) return FRSDG_ERR_UNSUPPORTED;
            const int nd_e = static_cast<int>(c->nd / c->E);
            const long long nb = static_cast<long long>(nd_e) * nd_e * c->E;
            if (!s->B) FR_TRY(hipMalloc(&s->B, sizeof(double) * nb));
            if (!s->pv) FR_TRY(hipMalloc(&s->pv, sizeof(double) * s->N));
            if (!s->pjv) FR_TRY(hipMalloc(&s->pjv, sizeof(double) * s->N));
            if (!s->info) FR_TRY(hipMalloc(&s->info, sizeof(int) * c->E));
            mark(0);
            const int keep = transpose;
            transpose      = 0; // blocks are probed with forward products
            for (int col = 0; col < 2; ++col)
                for (int m = 0; m < nd_e; ++m)
                {
                    hipLaunchKernelGGL(k_probe, dim3(grid(s->N)), dim3(256), 0, st, c->E, c->d.nx, nd_e, m, col, s->pv);
                    int rc = jv(s->pv, s->pjv);
                    if (rc) return rc;
                    hipLaunchKernelGGL(k_probe_store, dim3(grid(s->N)), dim3(256), 0, st, c->E, c->d.nx, nd_e, m, col,
                                       s->pjv, s->B);
                }
            transpose = keep;
            mark(1);
            hipLaunchKernelGGL(k_block_inverse, dim3(static_cast<unsigned>(c->E)), dim3(256), 0, st, nd_e, s->B, s->info);
            if (!s->Bf) FR_TRY(hipMalloc(&s->Bf, sizeof(float) * nb));
            hipLaunchKernelGGL(k_to_float, dim3(grid(nb)), dim3(256), 0, st, nb, s->B, s->Bf);
            FR_TRY(hipGetLastError());
            mark(2);
            FR_TRY(hipEventSynchronize(s->ev[2]));
            stats->t_pc_assemble += between(0, 1);
            stats->t_pc_factor += between(1, 2);
            return FRSDG_OK;
        }

        // GMRES(m): solve J x = b (x returned in s->du), right preconditioned
        int gmres(const double* b, double* x, int& its)
        {
            const int m = std::max(1, o->ksp_restart);
            const long long N = s->N;
            std::vector<double> H((m + 1) * m), g(m + 1), cs(m), sn(m), y(m);
            FR_TRY(hipMemsetAsync(x, 0, sizeof(double) * N, st));
            double beta0 = -1.0;
            its = 0;
            // PETSc's GMRES default is left preconditioning with the preconditioned
            // residual norm (-ksp_pc_side left); right keeps the true residual norm;
            // flexible (FGMRES) keeps the preconditioned basis Z_j = M^-1 v_j and
            // forms the solution from it, so the residual estimate stays the true
            // residual even where M^-1 is applied with large rounding growth
            const bool left = o->ksp_pc_side == 1 && o->precond != 0;
            const bool flex = o->ksp_pc_side == 2 && o->precond != 0;
            if (flex && (!s->Z || s->mZ < m))
            {
                if (s->Z) FR_TRY(hipFree(s->Z));
                s->Z  = nullptr;
                s->mZ = 0;
                FR_TRY(hipMalloc(&s->Z, sizeof(double) * m * N));
                s->mZ = m;
            }
            // r = b - J x  (x = 0 on the first cycle)
            bool first = true;
            while (true)
            {
                double* r = s->V; // V_0 holds r then v_0
                if (left)
                {
                    // left preconditioning: r = M^-1 (b - A x), x is the solution itself
                    int rc;
                    if (first)
                        rc = prec(b, r);
                    else
                    {
                        if ((rc = jv(x, s->w))) return rc;
                        hipLaunchKernelGGL(k_waxpby, dim3(grid(N)), dim3(256), 0, st, N, 1.0, b, -1.0, s->w, s->z);
                        rc = prec(s->z, r);
                    }
                    if (rc) return rc;
                }
                else if (first)
                    FR_TRY(hipMemcpyAsync(r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
                else if (flex)
                {
                    // x is the solution itself
                    int rc = jv(x, s->w);
                    if (rc) return rc;
                    hipLaunchKernelGGL(k_waxpby, dim3(grid(N)), dim3(256), 0, st, N, 1.0, b, -1.0, s->w, r);
                }
                else
                {
                    int rc = prec(x, s->z);
                    if (rc) return rc;
                    // x holds the unpreconditioned coordinates; the solution is M^-1 x
                    if ((rc = jv(s->z, s->w))) return rc;
                    hipLaunchKernelGGL(k_waxpby, dim3(grid(N)), dim3(256), 0, st, N, 1.0, b, -1.0, s->w, r);
                }
                first = false;
                const double beta = norm(r);
                if (beta0 < 0)
                {
                    beta0 = beta;
                    if (hist) hist->push_back(beta);
                }
                if (!(beta > o->ksp_rtol * beta0) || beta == 0.0 || its >= o->ksp_max_it) break;
                hipLaunchKernelGGL(k_axpby, dim3(grid(N)), dim3(256), 0, st, N, 0.0, r, 1.0 / beta, r);
                std::fill(g.begin(), g.end(), 0.0);
                g[0] = beta;
                int j = 0;
                bool done = false;
                for (; j < m && its < o->ksp_max_it; ++j)
                {
                    ++its;
                    int rc;
                    mark(0);
                    if (left)
                    {
                        if ((rc = jv(s->V + j * N, s->z))) return rc;
                        mark(1);
                        if ((rc = prec(s->z, s->w))) return rc;
                    }
                    else
                    {
                        double* zj = flex ? s->Z + j * N : s->z;
                        if ((rc = prec(s->V + j * N, zj))) return rc;
                        mark(1);
                        if ((rc = jv(zj, s->w))) return rc;
                    }
                    mark(2);
                    // classical Gram-Schmidt: h = V^T w; w -= V h; ||w|| (optionally a second pass)
                    std::vector<double> hh(j + 1, 0.0), h2(j + 1);
                    double nw2 = 0.0;
                    for (int pass = 0; pass < (o->ksp_cgs_refine ? 2 : 1); ++pass)
                    {
                        if ((rc = mdot(s->V, j + 1, s->w))) return rc;
                        hipLaunchKernelGGL(k_maxpy_norm, dim3(s->nb), dim3(RB), 0, st, N, s->V, j + 1, s->red, s->w,
                                           s->nb, s->part);
                        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, st, s->nb, s->part, s->red + (m + 1));
                        FR_TRY(hipMemcpyAsync(h2.data(), s->red, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, st));
                        FR_TRY(hipMemcpyAsync(&nw2, s->red + (m + 1), sizeof(double), hipMemcpyDeviceToHost, st));
                        FR_TRY(hipStreamSynchronize(st));
                        for (int i = 0; i <= j; ++i) hh[i] += h2[i];
                    }
                    mark(3);
                    FR_TRY(hipEventSynchronize(s->ev[3]));
                    stats->t_pc_apply += left ? between(1, 2) : between(0, 1);
                    stats->t_jv += left ? between(0, 1) : between(1, 2);
                    stats->t_orth += between(2, 3);
                    const double hn = std::sqrt(nw2);
                    for (int i = 0; i <= j; ++i) H[i * m + j] = hh[i];
                    H[(j + 1) * m + j] = hn;
                    if (hn > 0.0)
                        hipLaunchKernelGGL(k_waxpby, dim3(grid(N)), dim3(256), 0, st, N, 1.0 / hn, s->w, 0.0, s->w,
                                           s->V + (j + 1) * N);
                    // Givens
                    for (int i = 0; i < j; ++i)
                    {
                        const double a = H[i * m + j], bb = H[(i + 1) * m + j];
                        H[i * m + j]       = cs[i] * a + sn[i] * bb;
                        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
                    }
                    const double a = H[j * m + j], bb = H[(j + 1) * m + j];
                    const double den = std::sqrt(a * a + bb * bb);
                    cs[j] = den > 0 ? a / den : 1.0;
                    sn[j] = den > 0 ? bb / den : 0.0;
                    H[j * m + j] = den;
                    H[(j + 1) * m + j] = 0.0;
                    g[j + 1] = -sn[j] * g[j];
                    g[j]     = cs[j] * g[j];
                    if (hist) hist->push_back(std::fabs(g[j + 1]));
                    if (std::fabs(g[j + 1]) <= o->ksp_rtol * beta0 || hn == 0.0)
                    {
                        ++j;
                        done = true;
                        break;
                    }
                }
                // y = H^-1 g ; x += V y
                for (int i = j - 1; i >= 0; --i)
                {
                    double sum = g[i];
                    for (int k2 = i + 1; k2 < j; ++k2) sum -= H[i * m + k2] * y[k2];
                    y[i] = H[i * m + i] != 0.0 ? sum / H[i * m + i] : 0.0;
                }
                FR_TRY(hipMemcpyAsync(s->red, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_mcomb, dim3(grid(N)), dim3(256), 0, st, N, flex ? s->Z : s->V, j, s->red, s->z);
                hipLaunchKernelGGL(k_axpby, dim3(grid(N)), dim3(256), 0, st, N, 1.0, s->z, 1.0, x);
                FR_TRY(hipStreamSynchronize(st));
                if (done || its >= o->ksp_max_it) break;
            }
            if (left || flex) return pc_status();
            // solution = M^-1 x
            int rc = prec(x, s->z);
            if (rc) return rc;
            FR_TRY(hipMemcpyAsync(x, s->z, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            return pc_status();
        }
        int pc_status() const { return o->precond >= 3 && !transpose ? ilu_df_status(*s->cp, st) : FRSDG_OK; }
    };

    int ensure_ws(frsdg_ctx* c, int m)
    {
        if (c->ws && c->ws->m >= m) return FRSDG_OK;
        free_solver(c);
        auto* s = new SolverWs{};
        c->ws   = s;
        s->N    = c->nd;
        s->m    = m;
        s->nb   = static_cast<int>(std::min<long long>(1024, (c->nd + RB - 1) / RB));
        for (hipEvent_t& e : s->ev) FR_TRY(hipEventCreate(&e));
        FR_TRY(hipMalloc(&s->V, sizeof(double) * (m + 1) * s->N));
        for (double** p : {&s->w, &s->z, &s->un, &s->ud, &s->F, &s->du, &s->ut}) FR_TRY(hipMalloc(p, sizeof(double) * s->N));
        FR_TRY(hipMalloc(&s->part, sizeof(double) * (m + 2) * s->nb));
        FR_TRY(hipMalloc(&s->red, sizeof(double) * (2 * m + 4)));
        return FRSDG_OK;
    }
} // namespace

extern "C" int frsdg_beuler_step(frsdg_ctx* c, double t, double dt, double* u, const frsdg_newton_opts* o,
                                 frsdg_newton_stats* stats, void* stream)
{
    if (!c || !u || !o || !(dt > 0.0)) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    if (o->precond < 0 || o->precond > 4 || o->jac_type < 0 || o->jac_type > 1 || o->ksp_restart < 1)
        return FRSDG_ERR_INVALID;
    if (o->precond >= 1 && o->precond <= 2 && c->d.dims != 2) return FRSDG_ERR_UNSUPPORTED;
    frsdg_newton_stats local{};
    frsdg_newton_stats* S = stats ? stats : &local;
    *S = frsdg_newton_stats{};
    int rc = ensure_ws(c, o->ksp_restart);
    if (rc) return rc;
    SolverWs* s = c->ws;
    hipStream_t st = static_cast<hipStream_t>(stream);
    FR_TRY(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    const long long N = s->N;
    const double sigma = 1.0 / dt;
    Ctx K{c, s, st, o, t + dt, sigma, u, S};
    auto g = K.grid(N);
    // F(x) = (x - u^n)/dt + R(x); with cond_sys the transient variable is
    // C(x) (TSSetTransientVariable): F(x) = (C(x) - C(u^n))/dt + R(C(x))
    const bool cs = c->d.cond_sys != 0;
    const unsigned gn = static_cast<unsigned>((c->nm.nodes + 255) / 256);
    if (cs)
    {
        if (!c->w1) FR_TRY(hipMalloc(&c->w1, sizeof(double) * c->nd));
        if (!c->w2) FR_TRY(hipMalloc(&c->w2, sizeof(double) * c->nd));
        hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, u, s->un);
    }
    else
        FR_TRY(hipMemcpyAsync(s->un, u, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    auto eval_F = [&](const double* x, double* F) -> int
    {
        S->res_evals++;
        if (!cs && c->d.dims == 2) return frsdg::ifunction_be(c, x, s->un, sigma, F, st);
        if (cs)
        {
            hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, x, c->w2);
            hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, c->w2, -sigma, s->un, s->ud);
        }
        else
            hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, x, -sigma, s->un, s->ud);
        return frsdg_ifunction(c, t + dt, x, s->ud, F, st);
    };
    if ((rc = eval_F(u, s->F))) return rc;
    double fn = K.norm(s->F);
    S->fnorm0 = fn;
    const double target = std::max(o->snes_rtol * fn, o->snes_atol);
    bool         step_conv = false;
    while (fn > target && S->newton_its < o->snes_max_it && std::isfinite(fn))
    {
        // the reference rebuilds its preconditioning matrix at every Newton
        // iteration (PETSc's default lag 1); pc_lag k > 1 rebuilds it every
        // k-th Newton iteration, counted across the steps of this context
        // (-snes_lag_preconditioner k -snes_lag_preconditioner_persists)
        if (o->precond >= 1)
        {
            const int  kind    = o->precond + (o->pc_fp32 ? 16 : 0); // the single-precision copies are part of it
            const bool rebuild = s->pc_built != kind || o->pc_lag <= 1 || s->pc_age >= o->pc_lag;
            if (rebuild)
            {
                s->pc_built = -1;
                if (o->precond >= 3)
                    rc = K.build_ilu(o->precond == 3 ? 0 : 2, eval_F);
                else
                    rc = K.build_block_jacobi();
                if (rc) return rc;
                s->pc_built = kind;
                s->pc_age   = 0;
                S->pc_builds++;
            }
            s->pc_age++;
        }
        // J du = -F
        hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, st, N, 0.0, s->F, -1.0, s->F);
        int its = 0;
        if ((rc = K.gmres(s->F, s->du, its))) return rc;
        S->ksp_its += its;
        // PETSc's default line search (SNESLINESEARCHBT, order 3): Armijo
        // test with alpha = 1e-4 on 0.5 ||F||^2, a quadratic then cubic
        // backtracking model, each new lambda clamped to [0.1, 0.5] of the
        // previous one, at most 40 cubic steps, minlambda 1e-12.  The initial
        // slope is F . (J du) from one more J*v product, made negative as
        // PETSc does (-|F . J du|, -1 when zero).
        double initslope = 0.0;
        {
            const int keep_t = K.transpose;
            K.transpose      = 0;
            rc               = K.jv(s->du, s->w);
            K.transpose      = keep_t;
            if (rc || (rc = K.mdot(s->F, 1, s->w))) return rc; // s->F holds -F here: only |.| is used
            FR_TRY(hipMemcpyAsync(&initslope, s->red, sizeof(double), hipMemcpyDeviceToHost, st));
            FR_TRY(hipStreamSynchronize(st));
        }
        initslope = -std::fabs(initslope);
        if (initslope == 0.0) initslope = -1.0;
        const double alpha = 1e-4, minlambda = 1e-12;
        auto try_step = [&](double lam, double& gnorm) -> int
        {
            hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, 1.0, u, lam, s->du, s->ut);
            int r2 = eval_F(s->ut, s->w);
            if (r2) return r2;
            gnorm = K.norm(s->w);
            if (!std::isfinite(gnorm)) gnorm = std::numeric_limits<double>::infinity();
            return FRSDG_OK;
        };
        auto armijo = [&](double lam, double gnorm)
        { return 0.5 * gnorm * gnorm <= 0.5 * fn * fn + lam * alpha * initslope; };
        double lam = 1.0, fnew = 0.0;
        bool   decreased = false;
        if ((rc = try_step(lam, fnew))) return rc;
        if (armijo(lam, fnew))
            decreased = true;
        else
        {
            // quadratic
            double lamprev = lam, gprev = fnew;
            double lt      = -initslope / (fnew * fnew - fn * fn - 2.0 * lam * initslope);
            if (!std::isfinite(lt)) lt = 0.1 * lam;
            if (lt > 0.5 * lam) lt = 0.5 * lam;
            lam = (lt <= 0.1 * lam) ? 0.1 * lam : lt;
            if ((rc = try_step(lam, fnew))) return rc;
            if (armijo(lam, fnew))
                decreased = true;
            // cubic
            for (int count = 0; !decreased && count < 40 && lam > minlambda; ++count)
            {
                const double t1 = 0.5 * (fnew * fnew - fn * fn) - lam * initslope;
                const double t2 = 0.5 * (gprev * gprev - fn * fn) - lamprev * initslope;
                const double a  = (t1 / (lam * lam) - t2 / (lamprev * lamprev)) / (lam - lamprev);
                const double b  = (-lamprev * t1 / (lam * lam) + lam * t2 / (lamprev * lamprev)) / (lam - lamprev);
                double       d  = b * b - 3.0 * a * initslope;
                if (d < 0.0) d = 0.0;
                lt = (a == 0.0) ? -initslope / (2.0 * b) : (-b + std::sqrt(d)) / (3.0 * a);
                if (!std::isfinite(lt)) lt = 0.1 * lam;
                lamprev = lam;
                gprev   = fnew;
                if (lt > 0.5 * lam) lt = 0.5 * lam;
                lam = (lt <= 0.1 * lam) ? 0.1 * lam : lt;
                if ((rc = try_step(lam, fnew))) return rc;
                if (armijo(lam, fnew)) decreased = true;
            }
        }
        if (!decreased) break; // line search failed (SNES_DIVERGED_LINE_SEARCH): keep u
        // step-norm test of SNESConvergedDefault: ||lambda du|| <= stol ||u_new||
        bool snorm_conv = false;
        if (o->snes_stol > 0.0)
        {
            const double ynorm = lam * K.norm(s->du);
            const double xnorm = K.norm(s->ut);
            snorm_conv = ynorm <= o->snes_stol * xnorm;
        }
        FR_TRY(hipMemcpyAsync(u, s->ut, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        FR_TRY(hipMemcpyAsync(s->F, s->w, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        fn = fnew;
        S->newton_its++;
        if (snorm_conv && fn > target)
        {
            step_conv = true; // SNES_CONVERGED_SNORM_RELATIVE
            break;
        }
    }
    S->fnorm     = fn;
    S->converged = fn <= target ? 1 : (step_conv ? 2 : 0);
    // With the filter or the WENO limiter on, the reference stabilises the
    // solution once per implicit step (TSSetPostStep, or TSSetPostEvaluate
    // under -do_post_ts false: petsc_solver_impl.hpp:582-596; both run at the
    // end of a one-stage step), on the state vector as it is (W under cond_sys).
    if ((rc = frsdg_apply_stabilization(c, u, stream))) return rc;
    FR_TRY(hipStreamSynchronize(st));
    S->seconds   = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return FRSDG_OK;
}

// Pseudo-transient continuation (TSPSEUDO with TSPseudoTimeStepDefault):
// backward-Euler steps with the switched-evolution-relaxation step size.
extern "C" int frsdg_pseudo_run(frsdg_ctx* c, double t, double dt0, double increment, int from_initial_dt,
                                double dt_max, double t_max, int max_steps, double* u,
                                const frsdg_newton_opts* o, double* dt_hist, double* fnorm_hist, double* time_hist,
                                int* steps, frsdg_newton_stats* stats, void* stream)
{
    if (!c || !u || !o || !steps || !(dt0 > 0.0) || !(increment > 0.0) || max_steps < 0) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    *steps = 0;
    int rc = ensure_ws(c, o->ksp_restart);
    if (rc) return rc;
    SolverWs*   s  = c->ws;
    s->pc_built    = -1; // a run starts with a fresh preconditioner (lagged ones persist within it)
    hipStream_t st = static_cast<hipStream_t>(stream);
    FR_TRY(hipStreamSynchronize(st));
    const auto t_start = std::chrono::steady_clock::now();
    const double t_end_tol = 1e-12 * (t_max > 0.0 ? t_max : 1.0);
    // f(u) = ||F(u, U_t = 0)||: the IFunction with a zero time derivative
    auto steady_norm = [&](double tt, double& f) -> int
    {
        if (hipMemsetAsync(s->ud, 0, sizeof(double) * s->N, st) != hipSuccess) return FRSDG_ERR_HIP;
        int r = frsdg_ifunction(c, tt, u, s->ud, s->w, stream);
        if (r) return r;
        frsdg_newton_stats ns{};
        Ctx K{c, s, st, o, tt, 0.0, u, &ns};
        f = K.norm(s->w);
        return FRSDG_OK;
    };
    // TSStep_Pseudo computes each step's size before the step, from the current
    // state (TSPseudoComputeTimeStep -> TSPseudoTimeStepDefault): the first
    // call records fnorm_initial = fnorm_previous = f(u_0), so the first step
    // already takes inc * dt0
    double f = 0.0;
    if ((rc = steady_norm(t, f))) return rc;
    const double f_init = f;
    frsdg_newton_stats tot{};
    tot.fnorm0 = f;
    double       f_prev = f, dt = dt0;
    for (int k = 0; k < max_steps; ++k)
    {
        if (t_max > 0.0 && t >= t_max - t_end_tol) break;
        double nd;
        if (f == 0.0)
            nd = 1.e12 * increment * dt;
        else if (from_initial_dt)
            nd = increment * dt0 * f_init / f;
        else
            nd = increment * dt * f_prev / f;
        if (dt_max > 0.0) nd = std::min(nd, dt_max);
        f_prev = f;
        dt     = nd;
        if (t_max > 0.0 && t + dt > t_max) dt = t_max - t; // -ts_exact_final_time matchstep
        frsdg_newton_stats ns{};
        if ((rc = frsdg_beuler_step(c, t, dt, u, o, &ns, stream))) return rc;
        tot.newton_its += ns.newton_its;
        tot.ksp_its += ns.ksp_its;
        tot.jv_evals += ns.jv_evals;
        tot.res_evals += ns.res_evals;
        tot.pc_builds += ns.pc_builds;
        tot.seconds += ns.seconds;
        tot.t_pc_assemble += ns.t_pc_assemble;
        tot.t_pc_factor += ns.t_pc_factor;
        tot.t_pc_apply += ns.t_pc_apply;
        tot.t_jv += ns.t_jv;
        tot.t_orth += ns.t_orth;
        tot.converged = ns.converged;
        t += dt;
        if ((rc = steady_norm(t, f))) return rc;
        tot.fnorm = f;
        if (dt_hist) dt_hist[k] = dt;
        if (fnorm_hist) fnorm_hist[k] = f;
        if (time_hist) time_hist[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        *steps = k + 1;
    }
    FR_TRY(hipStreamSynchronize(st));
    if (stats) *stats = tot;
    return FRSDG_OK;
}

// Adjoint linear solve (sigma I + dR/du(u))^T lambda = rhs by GMRES with the
// adjoint kernel as operator (and the transposed block-Jacobi preconditioner).
extern "C" int frsdg_adjoint_solve(frsdg_ctx* c, double t, double sigma, const double* u, const double* rhs,
                                   double* lambda, const frsdg_newton_opts* o, frsdg_newton_stats* stats,
                                   void* stream)
{
    if (!c || !u || !rhs || !lambda || !o) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(rhs) || !aligned(lambda)) return FRSDG_ERR_ALIGN;
    if (o->precond < 0 || o->precond > 4) return FRSDG_ERR_INVALID;
    if (o->precond >= 1 && o->precond <= 2 && c->d.dims != 2) return FRSDG_ERR_UNSUPPORTED;
    frsdg_newton_stats local{};
    frsdg_newton_stats* S = stats ? stats : &local;
    *S = frsdg_newton_stats{};
    int rc = ensure_ws(c, o->ksp_restart);
    if (rc) return rc;
    SolverWs* s = c->ws;
    s->pc_built = -1; // this solve builds its own preconditioner into the shared buffers
    hipStream_t st = static_cast<hipStream_t>(stream);
    FR_TRY(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    Ctx K{c, s, st, o, t, sigma, u, S};
    if (o->precond >= 3)
    {
        // the coloured matrix of F(x) = IFunction(t, x, sigma (T(x) - T(u))) at u;
        // its ILU(0) applied transposed
        const long long N = s->N;
        const bool cs = c->d.cond_sys != 0;
        const unsigned gn = static_cast<unsigned>((c->nm.nodes + 255) / 256);
        const unsigned g  = static_cast<unsigned>((N + 255) / 256);
        if (cs)
        {
            if (!c->w2) FR_TRY(hipMalloc(&c->w2, sizeof(double) * c->nd));
            hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, u, s->un);
        }
        else
            FR_TRY(hipMemcpyAsync(s->un, u, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        auto eval_F = [&](const double* xx, double* F) -> int
        {
            S->res_evals++;
            if (!cs && c->d.dims == 2) return frsdg::ifunction_be(c, xx, s->un, sigma, F, st);
            if (cs)
            {
                hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, xx, c->w2);
                hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, c->w2, -sigma, s->un, s->ud);
            }
            else
                hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, xx, -sigma, s->un, s->ud);
            return frsdg_ifunction(c, t, xx, s->ud, F, st);
        };
        if ((rc = K.build_ilu(o->precond == 3 ? 0 : 2, eval_F))) return rc;
    }
    else if (o->precond >= 1)
        if ((rc = K.build_block_jacobi())) return rc;
    K.transpose = 1;
    FR_TRY(hipMemcpyAsync(s->F, rhs, sizeof(double) * s->N, hipMemcpyDeviceToDevice, st));
    S->fnorm0 = K.norm(s->F);
    int its = 0;
    if ((rc = K.gmres(s->F, lambda, its))) return rc;
    S->ksp_its = its;
    // true residual of the adjoint system
    if ((rc = K.jv(lambda, s->w))) return rc;
    hipLaunchKernelGGL(k_waxpby, dim3(K.grid(s->N)), dim3(256), 0, st, s->N, 1.0, s->F, -1.0, s->w, s->w);
    S->fnorm     = K.norm(s->w);
    S->converged = S->fnorm <= o->ksp_rtol * S->fnorm0 * 10.0;
    S->seconds   = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return FRSDG_OK;
}

// The coloured preconditioning matrix at u for F(x) = IFunction(t, x,
// sigma (T(x) - T(u_n))) (frsdg_color.hpp).
extern "C" int frsdg_color_jacobian(frsdg_ctx* c, double t, const double* u, const double* u_n, double sigma,
                                    int method, double* P, void* stream)
{
    if (!c || !u || !P || method < 0 || method > 2) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(P) || (u_n && !aligned(u_n))) return FRSDG_ERR_ALIGN;
    int rc = ensure_ws(c, c->ws ? c->ws->m : 1);
    if (rc) return rc;
    SolverWs* s = c->ws;
    if ((rc = color_plan_ensure(c, s))) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const long long N = s->N;
    const bool cs = c->d.cond_sys != 0;
    const unsigned gn = static_cast<unsigned>((c->nm.nodes + 255) / 256);
    const unsigned g  = static_cast<unsigned>((N + 255) / 256);
    const double* base = u_n ? u_n : u;
    if (cs)
    {
        if (!c->w2) FR_TRY(hipMalloc(&c->w2, sizeof(double) * c->nd));
        hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, base, s->un);
    }
    else
        FR_TRY(hipMemcpyAsync(s->un, base, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    frsdg_newton_opts oo{};
    frsdg_newton_stats ns{};
    Ctx K{c, s, st, &oo, t, sigma, u, &ns};
    auto eval_F = [&](const double* x, double* F) -> int
    {
        if (!cs && c->d.dims == 2) return frsdg::ifunction_be(c, x, s->un, sigma, F, st);
        if (cs)
        {
            hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, x, c->w2);
            hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, c->w2, -sigma, s->un, s->ud);
        }
        else
            hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, x, -sigma, s->un, s->ud);
        return frsdg_ifunction(c, t, x, s->ud, F, st);
    };
    if ((rc = K.color_assemble(method, eval_F, P))) return rc;
    FR_TRY(hipStreamSynchronize(st));
    return FRSDG_OK;
}

extern "C" int frsdg_ilu0_factor(frsdg_ctx* c, double* P, double* Dinv, void* stream)
{
    if (!c || !P || !Dinv) return FRSDG_ERR_INVALID;
    if (!aligned(P) || !aligned(Dinv)) return FRSDG_ERR_ALIGN;
    int rc = ensure_ws(c, c->ws ? c->ws->m : 1);
    if (rc || (rc = color_plan_ensure(c, c->ws))) return rc;
    return ilu_factor(c->ws, static_cast<hipStream_t>(stream), P, Dinv);
}

extern "C" int frsdg_ilu0_apply(frsdg_ctx* c, const double* P, const double* Dinv, const double* x, double* z,
                                void* stream)
{
    if (!c || !P || !Dinv || !x || !z) return FRSDG_ERR_INVALID;
    if (!aligned(P) || !aligned(Dinv) || !aligned(x) || !aligned(z)) return FRSDG_ERR_ALIGN;
    int rc = ensure_ws(c, c->ws ? c->ws->m : 1);
    if (rc || (rc = color_plan_ensure(c, c->ws))) return rc;
    if ((rc = ilu_apply(c->ws, static_cast<hipStream_t>(stream), P, Dinv, x, z))) return rc;
    return ilu_df_status(*c->ws->cp, static_cast<hipStream_t>(stream));
}

// (sigma I + dR/du(u)) x = b by the device GMRES, with the residual history.
extern "C" int frsdg_linear_solve(frsdg_ctx* c, double t, double sigma, const double* u, const double* b, double* x,
                                  const frsdg_newton_opts* o, frsdg_newton_stats* stats, double* hist, int hist_cap,
                                  int* hist_len, void* stream)
{
    if (!c || !u || !b || !x || !o || hist_cap < 0 || (hist_cap > 0 && (!hist || !hist_len))) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(b) || !aligned(x)) return FRSDG_ERR_ALIGN;
    if (o->precond < 0 || o->precond > 4 || o->jac_type < 0 || o->jac_type > 1 || o->ksp_restart < 1)
        return FRSDG_ERR_INVALID;
    if (o->precond >= 1 && o->precond <= 2 && c->d.dims != 2) return FRSDG_ERR_UNSUPPORTED;
    frsdg_newton_stats local{};
    frsdg_newton_stats* S = stats ? stats : &local;
    *S = frsdg_newton_stats{};
    int rc = ensure_ws(c, o->ksp_restart);
    if (rc) return rc;
    SolverWs* s = c->ws;
    s->pc_built = -1; // this solve builds its own preconditioner into the shared buffers
    hipStream_t st = static_cast<hipStream_t>(stream);
    FR_TRY(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    const long long N = s->N;
    Ctx K{c, s, st, o, t, sigma, u, S};
    std::vector<double> h;
    K.hist = &h;
    if (o->precond >= 3)
    {
        // the preconditioning matrix of F(x) = IFunction(t, x, sigma (T(x) - T(u))) at u
        const bool cs = c->d.cond_sys != 0;
        const unsigned gn = static_cast<unsigned>((c->nm.nodes + 255) / 256);
        const unsigned g  = static_cast<unsigned>((N + 255) / 256);
        if (cs)
        {
            if (!c->w2) FR_TRY(hipMalloc(&c->w2, sizeof(double) * c->nd));
            hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, u, s->un);
        }
        else
            FR_TRY(hipMemcpyAsync(s->un, u, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        auto eval_F = [&](const double* xx, double* F) -> int
        {
            S->res_evals++;
            if (!cs && c->d.dims == 2) return frsdg::ifunction_be(c, xx, s->un, sigma, F, st);
            if (cs)
            {
                hipLaunchKernelGGL(k_p2c, dim3(gn), dim3(256), 0, st, c->nm, xx, c->w2);
                hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, c->w2, -sigma, s->un, s->ud);
            }
            else
                hipLaunchKernelGGL(k_waxpby, dim3(g), dim3(256), 0, st, N, sigma, xx, -sigma, s->un, s->ud);
            return frsdg_ifunction(c, t, xx, s->ud, F, st);
        };
        if ((rc = K.build_ilu(o->precond == 3 ? 0 : 2, eval_F))) return rc;
    }
    else if (o->precond >= 1)
        if ((rc = K.build_block_jacobi())) return rc;
    FR_TRY(hipMemcpyAsync(s->F, b, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    S->fnorm0 = K.norm(s->F);
    int its = 0;
    if ((rc = K.gmres(s->F, x, its))) return rc;
    S->ksp_its = its;
    if ((rc = K.jv(x, s->w))) return rc;
    hipLaunchKernelGGL(k_waxpby, dim3(K.grid(N)), dim3(256), 0, st, N, 1.0, s->F, -1.0, s->w, s->w);
    S->fnorm     = K.norm(s->w);
    S->converged = S->fnorm <= o->ksp_rtol * S->fnorm0 * 10.0;
    S->seconds   = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (hist_len)
    {
        const int n = static_cast<int>(std::min<size_t>(h.size(), static_cast<size_t>(hist_cap)));
        for (int i = 0; i < n; ++i) hist[i] = h[i];
        *hist_len = n;
    }
    return FRSDG_OK;
}

extern "C" int frsdg_ilu0_apply_transpose(frsdg_ctx* c, const double* P, const double* Dinv, const double* x, double* z,
                                          void* stream)
{
    if (!c || !P || !Dinv || !x || !z) return FRSDG_ERR_INVALID;
    if (!aligned(P) || !aligned(Dinv) || !aligned(x) || !aligned(z)) return FRSDG_ERR_ALIGN;
    int rc = ensure_ws(c, c->ws ? c->ws->m : 1);
    if (rc || (rc = color_plan_ensure(c, c->ws))) return rc;
    return ilu_apply_T(c->ws, static_cast<hipStream_t>(stream), P, Dinv, x, z);
}

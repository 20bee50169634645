// frsdg_limiter.hpp — element stabilisation by the WENO limiter with shock
// sensor (filter.hpp:486-679, dispatch :714-755) inside limit_solution
// (physics_impl_2D.hpp:76-124, physics_impl_1D.hpp:96-121).
//
// Per line of an element (2D: rows along xi with the left/right neighbours,
// columns along eta with the bottom/top neighbours; 1D: the element's line
// per equation):
//   beta  = sum_{s=1}^{n-1} 4^s * 1/2 sum_i w_i (D^s p)_i^2     (smoothness)
//   om_k  = d_k / (1e-6 + beta_k)^2, normalised                 (d = linear weights)
//   shock = sum_k |om_k / d_k - 1|^theta / (|1/d_min - 1|^theta + 2) >= threshold
//   shock: out = om_0 (p_prev - avg_prev + avg) + om_1 p + om_2 (p_next - avg_next + avg)
//   else:  out = p;         2D result = (row result + column result) / 2.
// Beyond a non-periodic boundary the neighbour is the reference's zero ghost
// layer (a zero line).  The spectral filter, when active, replaces this
// output entirely (it is applied by the filter kernels instead).
//
// Arithmetic order is the reference's: every sum left to right from 0.0 and
// no fused multiply-add (its x86 build has none).  The smoothness indicator
// applies D up to n-1 times and amplifies rounding, so this order decides
// which lines the sensor flags; with it the limiter matches the CPU
// restatement bit for bit.
//
// 2D runs as two kernels: k_weno_beta2d (one thread per element line and
// equation, beta into a [E][2][n][4] buffer) and k_weno_apply2d (a workgroup
// of EPB elements: line tasks read the three betas, limit or copy their line
// into LDS, then each node writes (row + column) / 2).  The output must not
// alias the input (neighbour lines are read while other workgroups write).
#pragma once
#include <hip/hip_runtime.h>

namespace frsdg
{
    struct LimP
    {
        double    D[81]; // n x n differentiation matrix, row-major
        double    w[9];  // GLL weights
        double    ideal[3];
        double    theta, threshold, inv_den;
        int       nx, ny, per_x, per_y;
        long long E;
    };

    template <int N>
    __device__ __forceinline__ double lim_smoothness(const LimP& P, const double (&p)[N])
    {
#pragma clang fp contract(off)
        double cur[N];
#pragma unroll
        for (int i = 0; i < N; ++i) cur[i] = p[i];
        double beta = 0.0, factor = 4.0;
#pragma unroll
        for (int s = 1; s <= N - 1; ++s)
        {
            double nxt[N];
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                double a = 0.0;
#pragma unroll
                for (int c = 0; c < N; ++c) a = a + P.D[r * N + c] * cur[c];
                nxt[r] = a;
            }
            double I = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
            {
                cur[i] = nxt[i];
                I      = I + P.w[i] * cur[i] * cur[i];
            }
            beta   = beta + factor * (0.5 * I);
            factor = factor * 4.0;
        }
        return beta;
    }

    template <int N>
    __device__ __forceinline__ double lim_average(const LimP& P, const double (&p)[N])
    {
#pragma clang fp contract(off)
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) s = s + P.w[i] * p[i];
        return 0.5 * s;
    }

    // ShockSensor::detect_shock (filter.hpp:643-676)
    __device__ __forceinline__ bool lim_sensor(const LimP& P, double b0, double b1, double b2, double om[3])
    {
#pragma clang fp contract(off)
        const double eps = 1.0e-6;
        const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
        om[0]              = P.ideal[0] / (e0 * e0);
        om[1]              = P.ideal[1] / (e1 * e1);
        om[2]              = P.ideal[2] / (e2 * e2);
        const double w_sum = ((0.0 + om[0]) + om[1]) + om[2];
        double       num   = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            om[k]          = om[k] / w_sum;
            const double v = ::fabs(om[k] / P.ideal[k] - 1.0);
            num            = num + (P.theta == 1.0 ? v : ::pow(v, P.theta));
        }
        return ::fabs(num * P.inv_den) >= P.threshold;
    }

    // WENOLimiter::apply_weno_scalar (filter.hpp:518-545) on one line
    template <int N>
    __device__ __forceinline__ void lim_weno(const LimP& P, const double (&pp)[N], const double (&pc)[N],
                                             const double (&pn)[N], const double om[3], double (&out)[N])
    {
#pragma clang fp contract(off)
        const double ap = lim_average<N>(P, pp), ac = lim_average<N>(P, pc), an = lim_average<N>(P, pn);
#pragma unroll
        for (int i = 0; i < N; ++i)
        {
            const double tp = pp[i] - ap + ac;
            const double tn = pn[i] - an + ac;
            out[i]          = om[0] * tp + om[1] * pc[i] + om[2] * tn;
        }
    }

    // neighbour element of e along dir (0: x, 1: y) at offset dd (-1/+1);
    // -1 = zero ghost beyond a non-periodic boundary
    __device__ __forceinline__ long long lim_neighbour(const LimP& P, long long e, int dir, int dd)
    {
        int i = static_cast<int>(e % P.nx), j = static_cast<int>(e / P.nx);
        if (dir == 0)
        {
            i += dd;
            if (i < 0 || i >= P.nx)
            {
                if (!P.per_x) return -1;
                i = i < 0 ? P.nx - 1 : 0;
            }
        }
        else
        {
            j += dd;
            if (j < 0 || j >= P.ny)
            {
                if (!P.per_y) return -1;
                j = j < 0 ? P.ny - 1 : 0;
            }
        }
        return static_cast<long long>(j) * P.nx + i;
    }

    template <int N>
    __device__ __forceinline__ void lim_load_line(const double* __restrict__ u, long long e, int dir, int ln, int q,
                                                  double (&v)[N])
    {
        constexpr int NN = N * N;
        if (e < 0)
        {
#pragma unroll
            for (int m = 0; m < N; ++m) v[m] = 0.0;
            return;
        }
        const double* b = u + e * NN * 4 + q;
#pragma unroll
        for (int m = 0; m < N; ++m) v[m] = dir == 0 ? b[(ln * N + m) * 4] : b[(m * N + ln) * 4];
    }

    // smoothness of every (element, direction, line, equation): [E][2][N][4]
    template <int N>
    __global__ __launch_bounds__(256) void k_weno_beta2d(const LimP P, const double* __restrict__ u,
                                                         double* __restrict__ beta)
    {
        const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (t >= P.E * 8 * N) return;
        const int       q   = static_cast<int>(t & 3);
        const long long r   = t >> 2;
        const int       ln  = static_cast<int>(r % N);
        const long long r2  = r / N;
        const int       dir = static_cast<int>(r2 & 1);
        const long long e   = r2 >> 1;
        double          v[N];
        lim_load_line<N>(u, e, dir, ln, q, v);
        beta[t] = lim_smoothness<N>(P, v);
    }

    template <int N>
    struct LimTile
    {
        static constexpr int TASKS = 8 * N;                           // line tasks per element
        static constexpr int EPB   = 256 / TASKS > 0 ? 256 / TASKS : 1; // elements per workgroup
    };

    template <int N>
    __global__ __launch_bounds__(256) void k_weno_apply2d(const LimP P, const double* __restrict__ u,
                                                          const double* __restrict__ beta, double* __restrict__ out)
    {
        constexpr int NN  = N * N;
        constexpr int EPB = LimTile<N>::EPB;
        __shared__ double X[EPB][NN * 4], Y[EPB][NN * 4];
        const int       tid = threadIdx.x;
        const long long e0  = static_cast<long long>(blockIdx.x) * EPB;
        if (tid < EPB * LimTile<N>::TASKS)
        {
            const int       q   = tid & 3;
            const int       r   = tid >> 2;
            const int       ln  = r % N;
            const int       dir = (r / N) & 1;
            const int       el  = r / (2 * N);
            const long long e   = e0 + el;
            if (e < P.E)
            {
                const long long ep = lim_neighbour(P, e, dir, -1), en = lim_neighbour(P, e, dir, 1);
                const long long bo = (dir * N + ln) * 4 + q;
                const double    bp = ep < 0 ? 0.0 : beta[ep * 8 * N + bo];
                const double    bc = beta[e * 8 * N + bo];
                const double    bn = en < 0 ? 0.0 : beta[en * 8 * N + bo];
                double          pc[N], res[N], om[3];
                lim_load_line<N>(u, e, dir, ln, q, pc);
                if (lim_sensor(P, bp, bc, bn, om))
                {
                    double pp[N], pn[N];
                    lim_load_line<N>(u, ep, dir, ln, q, pp);
                    lim_load_line<N>(u, en, dir, ln, q, pn);
                    lim_weno<N>(P, pp, pc, pn, om, res);
                }
                else
                {
#pragma unroll
                    for (int m = 0; m < N; ++m) res[m] = pc[m];
                }
#pragma unroll
                for (int m = 0; m < N; ++m)
                {
                    if (dir == 0)
                        X[el][(ln * N + m) * 4 + q] = res[m];
                    else
                        Y[el][(m * N + ln) * 4 + q] = res[m];
                }
            }
        }
        __syncthreads();
        for (int t = tid; t < EPB * NN * 4; t += 256)
        {
            const int       el = t / (NN * 4);
            const int       o  = t - el * NN * 4;
            const long long e  = e0 + el;
            if (e < P.E) out[e * NN * 4 + o] = 0.5 * (X[el][o] + Y[el][o]);
        }
    }

    // 1D: one thread per (element, equation); layout [E][3][N]
    template <int N>
    __global__ __launch_bounds__(64) void k_weno1d(const LimP P, const double* __restrict__ u,
                                                   double* __restrict__ out)
    {
        const long long t = static_cast<long long>(blockIdx.x) * 64 + threadIdx.x;
        if (t >= P.E * 3) return;
        const long long e = t / 3;
        const int       q = static_cast<int>(t - e * 3);
        long long       ep = e - 1, en = e + 1;
        if (ep < 0) ep = P.per_x ? P.E - 1 : -1;
        if (en >= P.E) en = P.per_x ? 0 : -1;
        double pp[N], pc[N], pn[N], res[N], om[3];
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
            pc[m] = u[(e * 3 + q) * N + m];
            pp[m] = ep < 0 ? 0.0 : u[(ep * 3 + q) * N + m];
            pn[m] = en < 0 ? 0.0 : u[(en * 3 + q) * N + m];
        }
        if (lim_sensor(P, lim_smoothness<N>(P, pp), lim_smoothness<N>(P, pc), lim_smoothness<N>(P, pn), om))
            lim_weno<N>(P, pp, pc, pn, om, res);
        else
        {
#pragma unroll
            for (int m = 0; m < N; ++m) res[m] = pc[m];
        }
#pragma unroll
        for (int m = 0; m < N; ++m) out[(e * 3 + q) * N + m] = res[m];
    }
} // namespace frsdg

// frsdg_peaks.hip — measured machine peaks that the benchmark reports beside
// the datasheet figures (libfrsdg_peaks.so; not part of the residual ABI).
//
//   frsdg_peak_fma   FP64 vector FMA rate: every thread runs 8 independent
//                    dependency chains of `iters` fused multiply-adds
//                    (16 FLOP per iteration), so issue, not latency, limits it;
//   frsdg_peak_copy  HBM stream: out[i] = in[i] over n doubles with 16-byte
//                    vector loads and stores, grid-stride.
#include <hip/hip_runtime.h>

namespace
{
    __global__ __launch_bounds__(256) void k_fma_loop(double* out, int iters, double a, double b)
    {
        const double s = 1.0 + 1e-9 * (blockIdx.x * 256 + threadIdx.x);
        double       x0 = s, x1 = s + 1, x2 = s + 2, x3 = s + 3, x4 = s + 4, x5 = s + 5, x6 = s + 6, x7 = s + 7;
        for (int i = 0; i < iters; ++i)
        {
            x0 = __builtin_fma(x0, a, b);
            x1 = __builtin_fma(x1, a, b);
            x2 = __builtin_fma(x2, a, b);
            x3 = __builtin_fma(x3, a, b);
            x4 = __builtin_fma(x4, a, b);
            x5 = __builtin_fma(x5, a, b);
            x6 = __builtin_fma(x6, a, b);
            x7 = __builtin_fma(x7, a, b);
        }
        // one store per thread keeps every chain live
        out[blockIdx.x * 256 + threadIdx.x] = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
    }

    __global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ in, double2* __restrict__ out,
                                                  long long n2)
    {
        const long long stride = static_cast<long long>(gridDim.x) * 256;
        for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n2; i += stride)
            out[i] = in[i];
    }
} // namespace

extern "C" {

// out: device array of blocks * 256 doubles.  FLOPs = 16 * iters * blocks * 256.
int frsdg_peak_fma(double* out, int blocks, int iters, void* stream)
{
    if (!out || blocks <= 0 || iters <= 0) return 1;
    hipLaunchKernelGGL(k_fma_loop, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), out, iters,
                       0.999999999, 1e-12);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// in/out: device arrays of n doubles (n even, 16-byte aligned).  Bytes moved = 16 n.
int frsdg_peak_copy(const double* in, double* out, long long n, int blocks, void* stream)
{
    if (!in || !out || n <= 0 || (n & 1) || blocks <= 0) return 1;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out), n / 2);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

} // extern "C"

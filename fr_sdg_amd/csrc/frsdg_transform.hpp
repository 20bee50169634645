// frsdg_transform.hpp — the positivity-preserving change of variables of the
// implicit path (-cond_sys; problem_def.hpp:36-70): the solver state holds
// (log rho, u[, v], log e) per node and the residual sees the conservative
// state C(W) = (rho, rho u[, rho v], rho (e + |u|^2 / 2)), rho = exp(log rho),
// e = exp(log e).
//
//   k_p2c        W -> C(W)                          (prim_to_cons_kernel)
//   k_c2p        C -> W                             (cons_to_prim_elem, impose_ic)
//   k_p2c_tan    (W, V) -> (C(W), dC/dW V)          (forward tangent of the chain)
//   k_p2c_cplx   (W, V, h) -> Re, Im/h of C(W + i h V)  (the reference's complex step,
//                petsc_solver_impl.hpp:297-326, 386-401)
//   k_p2c_vjp    (W, G) -> (dC/dW)^T G              (adjoint of the chain)
//
// Layouts: 2D nodes interleave their 4 variables ((l n + k) 4 + eq); 1D
// elements store [eq][node] blocks of n.  All kernels are pointwise (one
// thread per node) and HBM-bound; they have internal linkage, as both
// translation units of the library launch them.
#pragma once
#include "frsdg_scalar.hpp"

#include <hip/hip_runtime.h>

namespace frsdg
{
    // node-variable addressing: 2D -> (node * 4 + q); 1D -> (e * 3 + q) * n + m
    struct NodeMap
    {
        int       dims, n;
        long long nodes; // total nodes
        __device__ __forceinline__ long long at(long long node, int q) const
        {
            if (dims == 2) return node * 4 + q;
            const long long e = node / n;
            return (e * 3 + q) * n + (node - e * n);
        }
    };

    // prim_to_cons_elem (problem_def.hpp:36-50): W = (r~, u, v, e~) or (r~, u, e~)
    template <class S>
    __device__ __forceinline__ void p2c_node(int dims, const S* W, S* C)
    {
        const S rho = exp_(W[0]);
        if (dims == 2)
        {
            const S u = W[1], v = W[2];
            const S kin = 0.5 * ((Sc<S>::c(0.0) + u * u) + v * v);
            C[0] = rho;
            C[1] = rho * u;
            C[2] = rho * v;
            C[3] = rho * (exp_(W[3]) + kin);
        }
        else
        {
            const S u   = W[1];
            const S kin = 0.5 * (Sc<S>::c(0.0) + u * u);
            C[0]        = rho;
            C[1]        = rho * u;
            C[2]        = rho * (exp_(W[2]) + kin);
        }
    }

    static __global__ __launch_bounds__(256) void k_p2c(const NodeMap M, const double* __restrict__ W, double* __restrict__ C)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        double w[4], c[4];
        if (M.dims == 2)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = W[M.at(i, q)];
            p2c_node<double>(2, w, c);
#pragma unroll
            for (int q = 0; q < 4; ++q) C[M.at(i, q)] = c[q];
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 3; ++q) w[q] = W[M.at(i, q)];
            p2c_node<double>(1, w, c);
#pragma unroll
            for (int q = 0; q < 3; ++q) C[M.at(i, q)] = c[q];
        }
    }

    // cons_to_prim_elem (problem_def.hpp:55-70); may run in place (W == C):
    // each thread reads its node before writing it
    static __global__ __launch_bounds__(256) void k_c2p(const NodeMap M, const double* C, double* W)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        const int nv = M.dims == 2 ? 4 : 3;
        double    c[4] = {0.0, 0.0, 0.0, 0.0}, w[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < nv) c[q] = C[M.at(i, q)];
        const double inv = 1.0 / c[0];
        w[0]             = ::log(c[0]);
        if (M.dims == 2)
        {
            const double u = c[1] * inv, v = c[2] * inv;
            const double kin = 0.5 * ((0.0 + u * u) + v * v);
            w[1]             = u;
            w[2]             = v;
            w[3]             = ::log(c[3] * inv - kin);
        }
        else
        {
            const double u   = c[1] * inv;
            const double kin = 0.5 * (0.0 + u * u);
            w[1]             = u;
            w[2]             = ::log(c[2] * inv - kin);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < nv) W[M.at(i, q)] = w[q];
    }

    static __global__ __launch_bounds__(256) void k_p2c_tan(const NodeMap M, const double* __restrict__ W,
                                                     const double* __restrict__ V, double* __restrict__ C,
                                                     double* __restrict__ CV)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        // fixed trip counts keep the Dual arrays in registers
        Dual w[4], c[4];
        if (M.dims == 2)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = Dual{W[M.at(i, q)], V[M.at(i, q)]};
            p2c_node<Dual>(2, w, c);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                C[M.at(i, q)]  = c[q].v;
                CV[M.at(i, q)] = c[q].d;
            }
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 3; ++q) w[q] = Dual{W[M.at(i, q)], V[M.at(i, q)]};
            p2c_node<Dual>(1, w, c);
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                C[M.at(i, q)]  = c[q].v;
                CV[M.at(i, q)] = c[q].d;
            }
        }
    }

    // complex state C(W + i h V) handed to the complex-step residual as
    // (Re, Im / h), which it recombines as Re + i h (Im / h)
    static __global__ __launch_bounds__(256) void k_p2c_cplx(const NodeMap M, const double* __restrict__ W,
                                                      const double* __restrict__ V, double h, double inv_h,
                                                      double* __restrict__ Cre, double* __restrict__ Cim_h)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        Cplx w[4], c[4];
        if (M.dims == 2)
        {
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = Cplx{W[M.at(i, q)], V[M.at(i, q)] * h};
            p2c_node<Cplx>(2, w, c);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                Cre[M.at(i, q)]   = c[q].re;
                Cim_h[M.at(i, q)] = c[q].im * inv_h;
            }
        }
        else
        {
#pragma unroll
            for (int q = 0; q < 3; ++q) w[q] = Cplx{W[M.at(i, q)], V[M.at(i, q)] * h};
            p2c_node<Cplx>(1, w, c);
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                Cre[M.at(i, q)]   = c[q].re;
                Cim_h[M.at(i, q)] = c[q].im * inv_h;
            }
        }
    }

    // (dC/dW)^T G at each node
    static __global__ __launch_bounds__(256) void k_p2c_vjp(const NodeMap M, const double* __restrict__ W,
                                                     const double* __restrict__ G, double* __restrict__ out)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        const double rho = ::exp(W[M.at(i, 0)]);
        if (M.dims == 2)
        {
            const double u = W[M.at(i, 1)], v = W[M.at(i, 2)], e = ::exp(W[M.at(i, 3)]);
            const double g0 = G[M.at(i, 0)], g1 = G[M.at(i, 1)], g2 = G[M.at(i, 2)], g3 = G[M.at(i, 3)];
            const double E = e + 0.5 * (u * u + v * v);
            out[M.at(i, 0)] = rho * (g0 + u * g1 + v * g2 + E * g3);
            out[M.at(i, 1)] = rho * (g1 + u * g3);
            out[M.at(i, 2)] = rho * (g2 + v * g3);
            out[M.at(i, 3)] = rho * e * g3;
        }
        else
        {
            const double u = W[M.at(i, 1)], e = ::exp(W[M.at(i, 2)]);
            const double g0 = G[M.at(i, 0)], g1 = G[M.at(i, 1)], g2 = G[M.at(i, 2)];
            const double E = e + 0.5 * u * u;
            out[M.at(i, 0)] = rho * (g0 + u * g1 + E * g2);
            out[M.at(i, 1)] = rho * (g1 + u * g2);
            out[M.at(i, 2)] = rho * e * g2;
        }
    }
} // namespace frsdg

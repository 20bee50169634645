// frsdg_capi.hip — context, setup kernels and the extern "C" entry points of
// libfrsdg.so (declared in include/frsdg.h).
#include "frsdg_adjoint.hpp"
#include "frsdg_ctx.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>

#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

using namespace frsdg;

namespace
{
    // ---------------- setup kernels ----------------
    // per-node composite metrics (physics_evaluator.hpp:235-313)
    __global__ void k_geom_vol(GeomP g, int nx, int ny, int n, const double* __restrict__ csi,
                               long long E, double* __restrict__ vmet)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        const long long NN = static_cast<long long>(n) * n;
        if (t >= E * NN) return;
        const long long e = t / NN;
        const int node = static_cast<int>(t - e * NN), k = node % n, l = node / n;
        const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        const double hx = 1.0 / static_cast<double>(nx), hy = 1.0 / static_cast<double>(ny);
        const double cx = hx * (static_cast<double>(i) + 0.5 * (1.0 + csi[k]));
        const double cy = hy * (static_cast<double>(j) + 0.5 * (1.0 + csi[l]));
        const VolM m    = composite_metrics(g, nx, ny, cx, cy);
        vmet[t * 4 + 0] = m.xi_x;
        vmet[t * 4 + 1] = m.xi_y;
        vmet[t * 4 + 2] = m.eta_x;
        vmet[t * 4 + 3] = m.eta_y;
    }
    // face normal (unit contravariant direction) and ||grad xi|| / ||grad eta||
    // at face points (physics_impl_2D.hpp:697-719, 805-819)
    __global__ void k_geom_face(GeomP g, int nx, int ny, int n, const double* __restrict__ csi,
                                int xdir, double* __restrict__ fm)
    {
        const long long t  = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        const long long nf = xdir ? static_cast<long long>(ny) * (nx + 1) : static_cast<long long>(ny + 1) * nx;
        if (t >= nf * n) return;
        const long long f = t / n;
        const int idx     = static_cast<int>(t - f * n);
        const double hx = 1.0 / static_cast<double>(nx), hy = 1.0 / static_cast<double>(ny);
        double cx, cy;
        if (xdir)
        {
            const int I = static_cast<int>(f % (nx + 1)), j = static_cast<int>(f / (nx + 1));
            cx = hx * (static_cast<double>(I) + 0.5 * (1.0 + csi[0]));
            cy = hy * (static_cast<double>(j) + 0.5 * (1.0 + csi[idx]));
        }
        else
        {
            const int i = static_cast<int>(f % nx), J = static_cast<int>(f / nx);
            cx = hx * (static_cast<double>(i) + 0.5 * (1.0 + csi[idx]));
            cy = hy * (static_cast<double>(J) + 0.5 * (1.0 + csi[0]));
        }
        const VolM m = composite_metrics(g, nx, ny, cx, cy);
        const Jac2 J = geom_jac(g, cx, cy);
        double a, b, sc;
        if (xdir)
        {
            a  = J.y_e;
            b  = 0.0 - J.x_e;
            sc = ::sqrt(m.xi_x * m.xi_x + m.xi_y * m.xi_y);
        }
        else
        {
            a  = 0.0 - J.y_c;
            b  = J.x_c;
            sc = ::sqrt(m.eta_x * m.eta_x + m.eta_y * m.eta_y);
        }
        const double inv = 1.0 / ::sqrt(a * a + b * b);
        fm[t * 4 + 0] = a * inv;
        fm[t * 4 + 1] = b * inv;
        fm[t * 4 + 2] = sc;
        fm[t * 4 + 3] = 0.0;
    }

    // quadrature weight of every DOF: w_k w_l detJ (2D, compute_error's
    // physics_impl_2D.hpp:253) or (h/2) w_k (1D, physics_impl_1D.hpp:260-312)
    __global__ void k_mass2d(GeomP g, int nx, int ny, int n, const double* __restrict__ csi,
                             const double* __restrict__ w, long long E, double* __restrict__ m)
    {
        const long long t  = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        const long long NN = static_cast<long long>(n) * n;
        if (t >= E * NN) return;
        const long long e = t / NN;
        const int node = static_cast<int>(t - e * NN), k = node % n, l = node / n;
        const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        const double cx = (1.0 / static_cast<double>(nx)) * (static_cast<double>(i) + 0.5 * (1.0 + csi[k]));
        const double cy = (1.0 / static_cast<double>(ny)) * (static_cast<double>(j) + 0.5 * (1.0 + csi[l]));
        const double v  = w[k] * w[l] * composite_metrics(g, nx, ny, cx, cy).detJ;
        for (int q = 0; q < 4; ++q) m[t * 4 + q] = v;
    }
    __global__ void k_mass1d(int ne, int n, double h, const double* __restrict__ w, double* __restrict__ m)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= static_cast<long long>(ne) * 3 * n) return;
        m[t] = 0.5 * h * w[t % n];
    }
    // boundary quadrature of face direction dir (0 L, 1 R, 2 T, 3 B) per face
    // node: (w_idx ds nx, w_idx ds ny), ds = |d(x,y)/d(face coordinate)|
    __global__ void k_wall2d(GeomP g, int nx, int ny, int n, const double* __restrict__ csi,
                             const double* __restrict__ w, int dir, double* __restrict__ out)
    {
        const int  along = (dir < 2) ? ny : nx;
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= static_cast<long long>(along) * n) return;
        const int el = static_cast<int>(t / n), idx = static_cast<int>(t % n);
        const double hx = 1.0 / static_cast<double>(nx), hy = 1.0 / static_cast<double>(ny);
        double cx, cy;
        if (dir < 2)
        {
            cx = (dir == 0) ? 0.0 : 1.0;
            cy = hy * (static_cast<double>(el) + 0.5 * (1.0 + csi[idx]));
        }
        else
        {
            cx = hx * (static_cast<double>(el) + 0.5 * (1.0 + csi[idx]));
            cy = (dir == 3) ? 0.0 : 1.0;
        }
        const Jac2 J = geom_jac(g, cx, cy);
        double a, b, ds;
        if (dir < 2)
        {
            a  = J.y_e;
            b  = 0.0 - J.x_e;
            ds = 0.5 * hy * ::sqrt(J.x_e * J.x_e + J.y_e * J.y_e);
        }
        else
        {
            a  = 0.0 - J.y_c;
            b  = J.x_c;
            ds = 0.5 * hx * ::sqrt(J.x_c * J.x_c + J.y_c * J.y_c);
        }
        const double inv = 1.0 / ::sqrt(a * a + b * b);
        out[t * 2 + 0] = w[idx] * ds * a * inv;
        out[t * 2 + 1] = w[idx] * ds * b * inv;
    }

    // ---------------- initial conditions / exact solutions ----------------
    // factory.hpp:311-352 (isentropic vortex, gamma hard-wired 1.4),
    // factory.hpp:423-454 (circular explosion), uniform state.
    struct ICP
    {
        int    kind;
        double prm[8];
        double t;
    };
    __device__ double gas_rhoE(const GasP& g, double rho, double vel, double pres)
    {
        if (g.kind == 0) return pres / (g.gamma - 1.0) + 0.5 * rho * vel * vel; // gas_models.hpp:81-86
        const double cov = 1.0 - g.b * rho, arr = g.a * rho * rho;            // :144-157
        return ((pres + arr) * cov) / (g.gamma - 1.0) - arr + 0.5 * rho * vel * vel;
    }
    __device__ void ic_state(const GasP& g, const ICP& c, double x, double y, double st[4])
    {
        if (c.kind == 0)
        {
            const double rho = c.prm[0], u = c.prm[1], v = c.prm[2], p = c.prm[3];
            st[0] = rho;
            st[1] = rho * u;
            st[2] = rho * v;
            st[3] = gas_rhoE(g, rho, ::sqrt(u * u + v * v), p);
            return;
        }
        if (c.kind == 4)
        {
            // smooth pressure pulse at rest: rho = rho0 (1 + A g), p = p0 (1 + A g)^gamma,
            // g = exp(-|x - x_c|^2 / w^2)
            const double dx = x - c.prm[2], dy = y - c.prm[3];
            const double gg = ::exp(-(dx * dx + dy * dy) / (c.prm[1] * c.prm[1]));
            const double rho = c.prm[4] * (1.0 + c.prm[0] * gg);
            st[0] = rho;
            st[1] = 0.0;
            st[2] = 0.0;
            st[3] = gas_rhoE(g, rho, 0.0, c.prm[5] * ::pow(1.0 + c.prm[0] * gg, g.gamma));
            return;
        }
        if (c.kind == 1)
        {
            const double u_0 = c.prm[0], v_0 = c.prm[1], beta = c.prm[2], r_v = c.prm[3], L = c.prm[4];
            const double gamma = 1.4, pi = 3.141592653589793238462643383279502884;
            double xc = L / 2.0 + u_0 * c.t, yc = L / 2.0 + v_0 * c.t;
            xc = xc - L * ::floor(xc / L);
            yc = yc - L * ::floor(yc / L);
            const double dx = x - xc, dy = y - yc;
            const double f  = 1.0 - ((dx * dx + dy * dy) / (r_v * r_v));
            const double ef = ::exp(f);
            const double rho = ::pow(1.0 - (beta * beta * (gamma - 1.0) * (ef * ef)) / (16.0 * gamma * pi * pi),
                                     1.0 / (gamma - 1.0));
            const double u = u_0 - (beta * ef * dy) / (2.0 * pi * r_v);
            const double v = v_0 + (beta * ef * dx) / (2.0 * pi * r_v);
            const double p = ::pow(rho, gamma);
            st[0] = rho;
            st[1] = rho * u;
            st[2] = rho * v;
            st[3] = (p / (gamma - 1.0)) + 0.5 * rho * (u * u + v * v);
            return;
        }
        const double dx = x - c.prm[5], dy = y - c.prm[6];
        const double r  = ::sqrt(dx * dx + dy * dy);
        double       rho = (r < c.prm[4]) ? c.prm[0] : c.prm[2];
        double       p   = (r < c.prm[4]) ? c.prm[1] : c.prm[3];
        if (c.prm[7] > 0.0)
        {
            // jump smoothed over a tanh layer of width prm[7]
            const double s = 0.5 * (1.0 - ::tanh((r - c.prm[4]) / c.prm[7]));
            rho = c.prm[2] + (c.prm[0] - c.prm[2]) * s;
            p   = c.prm[3] + (c.prm[1] - c.prm[3]) * s;
        }
        st[0] = rho;
        st[1] = rho * 0.0;
        st[2] = rho * 0.0;
        st[3] = gas_rhoE(g, rho, 0.0, p);
    }
    // node position as impose_ic places it (physics_impl_2D.hpp:163-177: Cartesian
    // h*(i + (1+csi)/2)); mapped geometries use the physical coordinates.
    __device__ void node_xy(const GeomP& g, int cart_pos, int nx, int ny, double lx, double ly,
                            const double* csi, int i, int j, int k, int l, double& x, double& y)
    {
        if (cart_pos)
        {
            const double hx = lx / nx, hy = ly / ny;
            x = hx * (i + 0.5 * (1.0 + csi[k]));
            y = hy * (j + 0.5 * (1.0 + csi[l]));
            return;
        }
        const double cx = (1.0 / static_cast<double>(nx)) * (static_cast<double>(i) + 0.5 * (1.0 + csi[k]));
        const double cy = (1.0 / static_cast<double>(ny)) * (static_cast<double>(j) + 0.5 * (1.0 + csi[l]));
        geom_phys(g, cx, cy, x, y);
    }
    // 1/radius per node for the radial source terms: 2D r = y/2 at the mapped
    // node (get_x_y_pos, physics_impl_2D.hpp:308-310), 1D r = (e + (1+csi)/2) h
    // (get_r, physics_evaluator.hpp:148-155); 0 within 1e-8 of the axis
    __global__ void k_rinv2d(GeomP g, int nx, int ny, int n, const double* __restrict__ csi, long long E,
                             double* __restrict__ out)
    {
        const long long t  = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        const long long NN = static_cast<long long>(n) * n;
        if (t >= E * NN) return;
        const long long e = t / NN;
        const int node = static_cast<int>(t - e * NN), k = node % n, l = node / n;
        double x, y;
        node_xy(g, 0, nx, ny, g.Lx, g.Ly, csi, static_cast<int>(e % nx), static_cast<int>(e / nx), k, l, x, y);
        const double r = 0.5 * y;
        out[t]         = (r > 1e-8) ? 1.0 / r : 0.0;
    }
    __global__ void k_rinv1d(int ne, int n, double h, const double* __restrict__ csi, double* __restrict__ out)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= static_cast<long long>(ne) * n) return;
        const int    e = static_cast<int>(t / n), m = static_cast<int>(t % n);
        const double r = (e + 0.5 * (csi[m] + 1.0)) * h;
        out[t]         = (r > 1e-8) ? 1.0 / r : 0.0;
    }
    __global__ void k_ic2d(GeomP g, GasP gas, ICP c, int cart_pos, int nx, int ny, int n, double lx,
                           double ly, const double* __restrict__ csi, long long E, double* __restrict__ u)
    {
        const long long t  = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        const long long NN = static_cast<long long>(n) * n;
        if (t >= E * NN) return;
        const long long e = t / NN;
        const int node = static_cast<int>(t - e * NN), k = node % n, l = node / n;
        const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
        double x, y, st[4];
        node_xy(g, cart_pos, nx, ny, lx, ly, csi, i, j, k, l, x, y);
        ic_state(gas, c, x, y, st);
        for (int q = 0; q < 4; ++q) u[t * 4 + q] = st[q];
    }
    // physics_impl_1D.hpp:207-244 with factory.hpp:196-214 (Sod) or a uniform state
    __global__ void k_ic1d(GasP gas, ICP c, int ne, int n, double lx, const double* __restrict__ csi,
                           double* __restrict__ u)
    {
        const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
        if (t >= static_cast<long long>(ne) * n) return;
        const int e = static_cast<int>(t / n), q = static_cast<int>(t % n);
        const double h = lx / ne, pos = h * (e + 0.5 * (1.0 + csi[q]));
        double rho, vel, p;
        if (c.kind == 3)
        {
            const bool left = pos < lx / 2.0;
            rho = left ? c.prm[0] : c.prm[3];
            vel = left ? c.prm[1] : c.prm[4];
            p   = left ? c.prm[2] : c.prm[5];
        }
        else
        {
            rho = c.prm[0];
            vel = c.prm[1];
            p   = c.prm[3];
        }
        double* sp = u + static_cast<long long>(e) * 3 * n;
        sp[0 * n + q] = rho;
        sp[1 * n + q] = rho * vel;
        sp[2 * n + q] = gas_rhoE(gas, rho, vel, p);
    }
    // compute_error partial sums (physics_impl_2D.hpp:203-265): one 4-vector per
    // block, reduced in fixed order by k_sum_blocks (deterministic).  With
    // cond_sys the stored node state is (log rho, u, v, log e) and is converted
    // to conservative variables first (physics_impl_2D.hpp:229-240).
    __global__ __launch_bounds__(256) void k_l2_part(GeomP g, GasP gas, ICP c, int nx, int ny, int n,
                                                     const double* __restrict__ csi,
                                                     const double* __restrict__ w, long long E,
                                                     const double* __restrict__ u, int cond_sys, double* part)
    {
        __shared__ double s[256][4];
        const long long NN = static_cast<long long>(n) * n;
        const long long t  = blockIdx.x * 256LL + threadIdx.x;
        double acc[4] = {0, 0, 0, 0};
        if (t < E * NN)
        {
            const long long e = t / NN;
            const int node = static_cast<int>(t - e * NN), k = node % n, l = node / n;
            const int i = static_cast<int>(e % nx), j = static_cast<int>(e / nx);
            const double cx = (1.0 / static_cast<double>(nx)) * (static_cast<double>(i) + 0.5 * (1.0 + csi[k]));
            const double cy = (1.0 / static_cast<double>(ny)) * (static_cast<double>(j) + 0.5 * (1.0 + csi[l]));
            double x, y, st[4];
            geom_phys(g, cx, cy, x, y);
            const VolM m = composite_metrics(g, nx, ny, cx, cy);
            ic_state(gas, c, x, y, st);
            const double cv = w[k] * w[l] * m.detJ;
            double       un[4];
            for (int q = 0; q < 4; ++q) un[q] = u[t * 4 + q];
            if (cond_sys)
            {
                const double wv[4] = {un[0], un[1], un[2], un[3]};
                p2c_node<double>(2, wv, un);
            }
            for (int q = 0; q < 4; ++q)
            {
                const double dd = un[q] - st[q];
                acc[q]          = cv * dd * dd;
            }
        }
        for (int q = 0; q < 4; ++q) s[threadIdx.x][q] = acc[q];
        __syncthreads();
        for (int stp = 128; stp > 0; stp >>= 1)
        {
            if (threadIdx.x < stp)
                for (int q = 0; q < 4; ++q) s[threadIdx.x][q] += s[threadIdx.x + stp][q];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int q = 0; q < 4; ++q) part[blockIdx.x * 4LL + q] = s[0][q];
    }
    __global__ __launch_bounds__(256) void k_sum_blocks(long long nb, const double* part, double* out)
    {
        __shared__ double s[256][4];
        double acc[4] = {0, 0, 0, 0};
        for (long long b = threadIdx.x; b < nb; b += 256)
            for (int q = 0; q < 4; ++q) acc[q] += part[b * 4 + q];
        for (int q = 0; q < 4; ++q) s[threadIdx.x][q] = acc[q];
        __syncthreads();
        for (int stp = 128; stp > 0; stp >>= 1)
        {
            if (threadIdx.x < stp)
                for (int q = 0; q < 4; ++q) s[threadIdx.x][q] += s[threadIdx.x + stp][q];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int q = 0; q < 4; ++q) out[q] = s[0][q];
    }

    // ---------------- dispatch over n = p+1 ----------------
    template <template <int> class F, class... A>
    int by_n(int n, A&&... a)
    {
        switch (n)
        {
            case 2: return F<2>::run(a...);
            case 3: return F<3>::run(a...);
            case 4: return F<4>::run(a...);
            case 5: return F<5>::run(a...);
            case 6: return F<6>::run(a...);
            case 7: return F<7>::run(a...);
            case 8: return F<8>::run(a...);
            case 9: return F<9>::run(a...);
        }
        return FRSDG_ERR_UNSUPPORTED;
    }

    template <class IO>
    struct Launch
    {
        template <int N>
        struct R2
        {
            static int run(const frsdg_ctx* c, const IO& io, hipStream_t s)
            {
                using T            = Tile2D<N, typename IO::S>;
                const long long nb = tile_count(c->p2.nx, c->p2.ny, T::TX, T::TY);
                if (c->p2.curved)
                    hipLaunchKernelGGL((k_res2d<N, IO, true>), dim3(static_cast<unsigned>(nb)), dim3(T::BS), 0, s,
                                       c->p2, io);
                else
                    hipLaunchKernelGGL((k_res2d<N, IO, false>), dim3(static_cast<unsigned>(nb)), dim3(T::BS), 0, s,
                                       c->p2, io);
                return hip_rc(hipGetLastError());
            }
        };
        template <int N>
        struct R1
        {
            static int run(const frsdg_ctx* c, const IO& io, hipStream_t s)
            {
                const int nb = (c->p1.ne + 63) / 64;
                hipLaunchKernelGGL((k_res1d<N, IO>), dim3(nb), dim3(64), 0, s, c->p1, io);
                return hip_rc(hipGetLastError());
            }
        };
    };
    // 2D forward tangent (k_tan2d: the residual's tiles, forward-mode pairs
    // only in the node and face fluxes)
    template <int N>
    struct Tan2
    {
        static int run(const frsdg_ctx* c, const IOTan& io, hipStream_t s)
        {
            using T            = Tile2D<N, double>;
            const long long nb = tile_count(c->p2.nx, c->p2.ny, T::TX, T::TY);
            if (c->p2.curved)
                hipLaunchKernelGGL((k_tan2d<N, true>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->p2, io);
            else
                hipLaunchKernelGGL((k_tan2d<N, false>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->p2, io);
            return hip_rc(hipGetLastError());
        }
    };

    template <class IO>
    int launch_res(const frsdg_ctx* c, const IO& io, hipStream_t s)
    {
        if (c->d.dims == 2) return by_n<Launch<IO>::template R2>(c->n, c, io, s);
        return by_n<Launch<IO>::template R1>(c->n, c, io, s);
    }

    template <int N>
    struct Filt
    {
        static int run(const frsdg_ctx* c, const double* u, double* o, hipStream_t s, int transpose = 0)
        {
            if (c->d.dims == 2)
            {
                constexpr int EPB = (256 / (N * N)) > 0 ? 256 / (N * N) : 1;
                const long long nb = (c->E + EPB - 1) / EPB;
                hipLaunchKernelGGL((k_filter2d<N>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->E,
                                   c->d_tab, u, o, transpose);
            }
            else
            {
                const int nb = (c->p1.ne + 63) / 64;
                hipLaunchKernelGGL((k_filter1d<N>), dim3(nb), dim3(64), 0, s, c->p1.ne, c->d_tab, u, o, transpose);
            }
            return hip_rc(hipGetLastError());
        }
    };

    template <int N>
    struct Adj2
    {
        static int run(const frsdg_ctx* c, const double* u, const double* w, const AdjOut& ao, double* out,
                       hipStream_t s)
        {
            const long long nl = c->E * 2 * N;
            hipLaunchKernelGGL((k_adj_lift<N>), dim3(static_cast<unsigned>((nl + 255) / 256)), dim3(256), 0, s,
                               c->E, c->d_tab, w, c->wb);
            using T            = Tile2D<N, double>;
            const long long nb = tile_count(c->p2.nx, c->p2.ny, T::TX, T::TY);
            const double*   wb = c->wb;
            if (c->p2.curved)
                hipLaunchKernelGGL((k_adj2d_tile<N, true>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->p2,
                                   u, w, wb, ao, out);
            else
                hipLaunchKernelGGL((k_adj2d_tile<N, false>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->p2,
                                   u, w, wb, ao, out);
            return hip_rc(hipGetLastError());
        }
    };
    template <int N>
    struct Adj1
    {
        static int run(const frsdg_ctx* c, const double* u, const double* w, const AdjOut& ao, double* out,
                       hipStream_t s)
        {
            const int nb = (c->p1.ne + 63) / 64;
            hipLaunchKernelGGL((k_adj1d<N>), dim3(nb), dim3(64), 0, s, c->p1, u, w, ao, out);
            return hip_rc(hipGetLastError());
        }
    };

    // one SSP-RK stage (k_res2d_stage in 2D, the stage output policy of the
    // residual kernel in 1D); filt: apply the post-step spectral filter (2D)
    template <int N>
    struct Stage
    {
        static int run(const frsdg_ctx* c, const IOStage& io, int filt, hipStream_t s)
        {
            if (c->d.dims != 2) return Launch<IOStage>::template R1<N>::run(c, io, s);
            using T            = Tile2D<N, double>;
            const long long nb = tile_count(c->p2.nx, c->p2.ny, T::TX, T::TY);
            if (c->p2.curved)
                hipLaunchKernelGGL((k_res2d_stage<N, true, true>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                                   c->p2, io, filt);
            else if (filt)
                hipLaunchKernelGGL((k_res2d_stage<N, false, true>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                                   c->p2, io, filt);
            else
                hipLaunchKernelGGL((k_res2d_stage<N, false, false>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                                   c->p2, io, filt);
            return hip_rc(hipGetLastError());
        }
    };

    // WENO limiter with shock sensor (limit_solution with -apply_weno and no
    // spectral filter); o must not alias u
    template <int N>
    struct Weno
    {
        static int run(const frsdg_ctx* c, const double* u, double* o, hipStream_t s)
        {
            if (c->d.dims == 2)
            {
                const long long nl = c->E * 8 * N;
                hipLaunchKernelGGL((k_weno_beta2d<N>), dim3(static_cast<unsigned>((nl + 255) / 256)), dim3(256), 0, s,
                                   c->lim, u, c->wbeta);
                constexpr int   EPB = LimTile<N>::EPB;
                const long long nb  = (c->E + EPB - 1) / EPB;
                hipLaunchKernelGGL((k_weno_apply2d<N>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, c->lim, u,
                                   static_cast<const double*>(c->wbeta), o);
            }
            else
            {
                const long long nt = c->E * 3;
                hipLaunchKernelGGL((k_weno1d<N>), dim3(static_cast<unsigned>((nt + 63) / 64)), dim3(64), 0, s, c->lim,
                                   u, o);
            }
            return hip_rc(hipGetLastError());
        }
    };

    bool weno_only(const frsdg_ctx* c) { return c->d.apply_weno && !c->d.apply_filter; }

    unsigned node_grid(const frsdg_ctx* c) { return static_cast<unsigned>((c->nm.nodes + 255) / 256); }

    // Health flags of stored states (bit 1 non-finite, bit 2 rho <= 0, bit 4
    // internal energy <= 0), one thread per node; 2D and 1D layouts
    __global__ __launch_bounds__(256) void k_health(NodeMap M, const double* __restrict__ u, int* flag)
    {
        const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
        if (i >= M.nodes) return;
        double o[4];
        if (M.dims == 2)
            ld4d(u + i * 4, o);
        else
        {
            o[0] = u[M.at(i, 0)];
            o[1] = u[M.at(i, 1)];
            o[2] = 0.0;
            o[3] = u[M.at(i, 2)];
        }
        health_check(o, flag);
    }

    int ensure_work(frsdg_ctx* c)
    {
        if (!c->health)
        {
            FR_TRY(hipMalloc(&c->health, sizeof(int)));
            FR_TRY(hipMemset(c->health, 0, sizeof(int)));
        }
        if (!c->w1) FR_TRY(hipMalloc(&c->w1, sizeof(double) * c->nd));
        if (!c->w2) FR_TRY(hipMalloc(&c->w2, sizeof(double) * c->nd));
        if (weno_only(c) && c->d.dims == 2 && !c->wbeta)
            FR_TRY(hipMalloc(&c->wbeta, sizeof(double) * c->E * 8 * c->n));
        return FRSDG_OK;
    }

    // limit_solution: the spectral filter when active (it replaces the WENO
    // result, filter.hpp:740-745), else the WENO limiter, else a copy
    int do_filter(frsdg_ctx* c, const double* u, double* out, hipStream_t s)
    {
        if (c->d.apply_filter) return by_n<Filt>(c->n, c, u, out, s);
        if (c->d.apply_weno)
        {
            int rc = ensure_work(c);
            if (rc) return rc;
            if (out != u) return by_n<Weno>(c->n, c, u, out, s);
            if ((rc = by_n<Weno>(c->n, c, u, c->w1, s))) return rc;
            FR_TRY(hipMemcpyAsync(out, c->w1, sizeof(double) * c->nd, hipMemcpyDeviceToDevice, s));
            return FRSDG_OK;
        }
        if (out != u) FR_TRY(hipMemcpyAsync(out, u, sizeof(double) * c->nd, hipMemcpyDeviceToDevice, s));
        return FRSDG_OK;
    }

    // one explicit step, stages fused with the residual kernel
    int do_step(frsdg_ctx* c, double dt, double* u, int stages, hipStream_t s)
    {
        // 2D stage outputs are checked in the stage kernel's store (except the
        // one that feeds the WENO limiter, checked after the limiter); 1D states
        // after the step
        int* hf    = c->d.dims == 2 ? c->health : nullptr;
        auto stage = [&](const double* us, double a0, double b, double cdt, double* out)
        {
            IOStage io{us, u, out, a0, b, cdt, hf};
            return by_n<Stage>(c->n, c, io, 0, s);
        };
        auto check_u = [&]() -> int
        {
            if (c->d.dims == 2 && !weno_only(c)) return FRSDG_OK;
            hipLaunchKernelGGL(k_health, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, c->health);
            return hip_rc(hipGetLastError());
        };
        int rc;
        if (stages == 3)
        {
            if ((rc = stage(u, 0.0, 1.0, dt, c->w1))) return rc;
            if ((rc = stage(c->w1, 0.75, 0.25, dt, c->w2))) return rc;
            if (c->d.apply_filter && c->d.dims == 2)
            {
                // last stage fused with the post-step filter
                IOStage io{c->w2, u, u, 1.0 / 3.0, 2.0 / 3.0, dt, hf};
                return by_n<Stage>(c->n, c, io, 1, s);
            }
            if (weno_only(c))
            {
                // last stage into w1 (free after stage two), then limit into u
                IOStage io{c->w2, u, c->w1, 1.0 / 3.0, 2.0 / 3.0, dt};
                if ((rc = by_n<Stage>(c->n, c, io, 0, s))) return rc;
                if ((rc = by_n<Weno>(c->n, c, c->w1, u, s))) return rc;
                return check_u();
            }
            if ((rc = stage(c->w2, 1.0 / 3.0, 2.0 / 3.0, dt, u))) return rc;
        }
        else if (stages == 4)
        {
            const double h = 0.5 * dt;
            if ((rc = stage(u, 0.0, 1.0, h, c->w1))) return rc;
            if ((rc = stage(c->w1, 0.0, 1.0, h, c->w2))) return rc;
            if ((rc = stage(c->w2, 2.0 / 3.0, 1.0 / 3.0, h, c->w1))) return rc;
            // last stage reads w1 (neighbour traces) and writes u (w2 when the
            // WENO limiter follows)
            IOStage io{c->w1, nullptr, weno_only(c) ? c->w2 : u, 0.0, 1.0, h, weno_only(c) ? nullptr : hf};
            if ((rc = by_n<Stage>(c->n, c, io, 0, s))) return rc;
            if (weno_only(c))
            {
                if ((rc = by_n<Weno>(c->n, c, c->w2, u, s))) return rc;
                return check_u();
            }
        }
        else
            return FRSDG_ERR_INVALID;
        if (c->d.apply_filter && (rc = by_n<Filt>(c->n, c, u, u, s))) return rc;
        return check_u();
    }

    GeomP geom_of(const frsdg_desc& d)
    {
        GeomP g{};
        g.kind = d.geometry;
        g.Lx   = d.lx;
        g.Ly   = d.ly;
        for (int i = 0; i < 4; ++i) g.prm[i] = d.geom_param[i];
        return g;
    }
} // namespace

namespace
{
    // File-scope scratch of frsdg_l2_error_noalloc: 4 partial sums per block
    // of 256 nodes plus the final 4-vector.  Each GPU holds its own copy
    // until the process ends; host code reaches it only through the device
    // address that hipGetSymbolAddress returns for the current device.
    static __device__ __align__(256) double scratch[4 * (FRSDG_L2_SCRATCH_BLOCKS + 1)];

    constexpr int MAX_DEVICES = 64;
    std::mutex    g_scratch_mu;
    double*       g_scratch_ptr[MAX_DEVICES] = {};

    // device address of `scratch` on the current device, looked up once per device
    int scratch_device_ptr(double** out)
    {
        int dev = 0;
        FR_TRY(hipGetDevice(&dev));
        if (dev < 0 || dev >= MAX_DEVICES) return FRSDG_ERR_UNSUPPORTED;
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev])
        {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FRSDG_ERR_HIP;
            g_scratch_ptr[dev] = static_cast<double*>(p);
        }
        *out = g_scratch_ptr[dev];
        return FRSDG_OK;
    }

    // compute_error into a caller-provided reduction buffer red (4 * (nb + 1)
    // doubles: the result, then the per-block partials); synchronises s
    int l2_error_into(const frsdg_ctx* c, const ICP& ic, const double* u, double* red, double* out, hipStream_t s)
    {
        const int       n   = c->n;
        const long long tot = c->E * n * n;
        const long long nb  = (tot + 255) / 256;
        hipLaunchKernelGGL(k_l2_part, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, geom_of(c->d), c->p2.gas, ic,
                           c->d.nx, c->d.ny, n, c->d_csi, c->d_csi + n, c->E, u, c->d.cond_sys, red + 4);
        hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(256), 0, s, nb, red + 4, red);
        FR_TRY(hipGetLastError());
        double h[4];
        FR_TRY(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, s));
        FR_TRY(hipStreamSynchronize(s));
        for (int q = 0; q < 4; ++q) out[q] = std::sqrt(h[q]);
        return FRSDG_OK;
    }
} // namespace

int frsdg::ifunction_be(frsdg_ctx* c, const double* x, const double* un, double sigma, double* F, hipStream_t s)
{
    if (c->d.dims != 2 || c->d.cond_sys) return FRSDG_ERR_UNSUPPORTED;
    IORes io{x, F, nullptr, 1.0};
    io.un    = un;
    io.shift = sigma;
    return launch_res(c, io, s);
}

extern "C" {

const char* frsdg_strerror(int code)
{
    switch (code)
    {
        case FRSDG_OK: return "success";
        case FRSDG_ERR_INVALID: return "invalid argument or descriptor";
        case FRSDG_ERR_HIP: return "HIP runtime error";
        case FRSDG_ERR_UNSUPPORTED: return "unsupported on this path";
        case FRSDG_ERR_ALIGN: return "device pointer not 16-byte aligned";
        case FRSDG_ERR_NUMERIC: return "numerical failure (zero pivot, or a non-finite / non-positive state)";
    }
    return "unknown error";
}

int frsdg_create(const frsdg_desc* d, frsdg_ctx** out)
{
    if (!d || !out) return FRSDG_ERR_INVALID;
    *out = nullptr;
    if (d->dims != 1 && d->dims != 2) return FRSDG_ERR_INVALID;
    if (d->p < 1 || d->p > 8 || d->nx < 1 || (d->dims == 2 && d->ny < 1)) return FRSDG_ERR_INVALID;
    if (d->gas_kind < 0 || d->gas_kind > 1 || d->riemann < 0 || d->riemann > 1) return FRSDG_ERR_INVALID;
    if (d->geometry < 0 || d->geometry > 4 || (d->dims == 1 && d->geometry != 0)) return FRSDG_ERR_INVALID;
    if (d->coord_sys < 0 || d->coord_sys > 1) return FRSDG_ERR_INVALID;
    for (int k = 0; k < (d->dims == 2 ? 4 : 2); ++k)
    {
        const int t = d->bc_type[k];
        if (!(t == 0 || t == 1 || t == 2 || t == 5 || (t == 6 && d->dims == 2))) return FRSDG_ERR_INVALID;
    }
    if (d->apply_filter && (d->filt_cutoff < 0 || d->filt_cutoff >= d->p)) return FRSDG_ERR_INVALID;
    // WENOLimiter ctor (filter.hpp:501-512) rejects gamma outside [0, 1]; 0 and 1
    // make a linear weight vanish and the sensor divide by it, so both are
    // rejected here as well
    if (d->apply_weno && !(d->weno_gamma > 0.0 && d->weno_gamma < 1.0 && d->shock_theta > 0.0))
        return FRSDG_ERR_INVALID;
    auto* c = new (std::nothrow) frsdg_ctx{};
    if (!c) return FRSDG_ERR_INVALID;
    c->d = *d;
    if (!make_tables(d->p, d->c_vcjh, d->filt_cutoff, d->filt_strength, d->filt_alpha, c->tb))
    {
        delete c;
        return FRSDG_ERR_INVALID;
    }
    const int n = c->n = d->p + 1;
    GasP gas{d->gas_kind, d->gamma, d->vdw_a, d->vdw_b};
    BCP  bc{};
    for (int k = 0; k < 4; ++k)
    {
        bc.type[k] = d->bc_type[k];
        for (int q = 0; q < 4; ++q) bc.state[k][q] = d->bc_state[k][q];
    }
    bc.freeze = d->tangent_bc_mode;
    std::vector<double> tab;
    tab.insert(tab.end(), c->tb.D.begin(), c->tb.D.end());
    tab.insert(tab.end(), c->tb.dL.begin(), c->tb.dL.end());
    tab.insert(tab.end(), c->tb.dR.begin(), c->tb.dR.end());
    tab.insert(tab.end(), c->tb.filt.begin(), c->tb.filt.end());
    // even-odd tables of D, of the filter and of D transposed (the adjoint's
    // lines) for the line kernels (eo_line, h = n / 2, c = h):
    // A[k][m] = (M[k][m] + M[n-1-k][m]) / 2, B[k][m] = (M[k][m] - M[n-1-k][m]) / 2
    // (k, m < h); odd n: M[k][c] (k < h), M[c][m] (m < h), M[c][c]
    {
        const int h = n / 2;
        std::vector<double> Dt(static_cast<size_t>(n) * n);
        for (int k = 0; k < n; ++k)
            for (int m = 0; m < n; ++m) Dt[k * n + m] = c->tb.D[m * n + k];
        for (const std::vector<double>* M : {&c->tb.D, &c->tb.filt, &Dt})
        {
            for (int sgn = 0; sgn < 2; ++sgn)
                for (int k = 0; k < h; ++k)
                    for (int m = 0; m < h; ++m)
                    {
                        const double x = (*M)[k * n + m], y = (*M)[(n - 1 - k) * n + m];
                        tab.push_back(0.5 * (sgn == 0 ? x + y : x - y));
                    }
            if (n % 2)
            {
                for (int k = 0; k < h; ++k) tab.push_back((*M)[k * n + h]);
                for (int m = 0; m < h; ++m) tab.push_back((*M)[h * n + m]);
                tab.push_back((*M)[h * n + h]);
            }
        }
    }
    int rc = FRSDG_OK;
    auto fail = [&](int code)
    {
        frsdg_destroy(c);
        return code;
    };
    if (hipMalloc(&c->d_tab, sizeof(double) * tab.size()) != hipSuccess) return fail(FRSDG_ERR_HIP);
    if (hipMemcpy(c->d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(FRSDG_ERR_HIP);
    if (hipMalloc(&c->d_csi, sizeof(double) * 2 * n) != hipSuccess) return fail(FRSDG_ERR_HIP);
    {
        std::vector<double> cw(c->tb.csi);
        cw.insert(cw.end(), c->tb.w.begin(), c->tb.w.end());
        if (hipMemcpy(c->d_csi, cw.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice) != hipSuccess)
            return fail(FRSDG_ERR_HIP);
    }
    {
        // WENO limiter constants (WENOLimiter / ShockSensor ctors, filter.hpp:501-512, :624-630)
        LimP& L = c->lim;
        for (int i = 0; i < n * n; ++i) L.D[i] = c->tb.D[i];
        for (int i = 0; i < n; ++i) L.w[i] = c->tb.w[i];
        const double gm    = d->weno_gamma;
        L.ideal[0]         = (1 - gm) * 0.5;
        L.ideal[1]         = gm;
        L.ideal[2]         = (1 - gm) * 0.5;
        const double d_min = std::min(std::min(L.ideal[0], L.ideal[1]), L.ideal[2]);
        L.theta            = d->shock_theta;
        L.threshold        = d->shock_threshold;
        L.inv_den          = 1.0 / (std::pow(std::abs(1.0 / d_min - 1.0), L.theta) + 3 - 1);
        L.nx               = d->nx;
        L.ny               = d->dims == 2 ? d->ny : 1;
        L.per_x            = d->bc_type[0] == 5 && d->bc_type[1] == 5;
        L.per_y            = d->dims == 2 && d->bc_type[2] == 5 && d->bc_type[3] == 5;
        L.E                = static_cast<long long>(L.nx) * L.ny;
    }
    c->nm.dims  = d->dims;
    c->nm.n     = n;
    c->nm.nodes = d->dims == 2 ? static_cast<long long>(d->nx) * d->ny * n * n : static_cast<long long>(d->nx) * n;
    if (d->dims == 1)
    {
        if (d->nx < 2 && !(d->bc_type[0] == 5 && d->bc_type[1] == 5)) return fail(FRSDG_ERR_INVALID);
        c->E  = d->nx;
        c->nd = 3LL * n * d->nx;
        c->p1 = Res1DP{d->nx, d->bc_type[0] == 5 && d->bc_type[1] == 5, d->riemann,
                       2.0 / (d->lx / d->nx), c->d_tab, gas, bc, nullptr};
        if (d->coord_sys == 1)
        {
            const long long tot = static_cast<long long>(d->nx) * n;
            if (hipMalloc(&c->d_rinv, sizeof(double) * tot) != hipSuccess) return fail(FRSDG_ERR_HIP);
            hipLaunchKernelGGL(k_rinv1d, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, nullptr, d->nx,
                               n, d->lx / d->nx, c->d_csi, c->d_rinv);
            if (hipDeviceSynchronize() != hipSuccess) return fail(FRSDG_ERR_HIP);
            c->p1.rinv = c->d_rinv;
        }
        *out  = c;
        return FRSDG_OK;
    }
    c->E  = static_cast<long long>(d->nx) * d->ny;
    if (c->E >= (1LL << 31) - 1024) return fail(FRSDG_ERR_UNSUPPORTED); // 32-bit element indices in kernels
    {
        // set-up of frsdg_l2_error_noalloc: the device address of its static
        // scratch on this device (once per device)
        double* scr = nullptr;
        if (scratch_device_ptr(&scr)) return fail(FRSDG_ERR_HIP);
    }
    c->nd = 4LL * n * n * c->E;
    Res2DP& P = c->p2;
    P.nx      = d->nx;
    P.ny      = d->ny;
    P.E       = c->E;
    P.per_x   = d->bc_type[0] == 5 && d->bc_type[1] == 5;
    P.per_y   = d->bc_type[2] == 5 && d->bc_type[3] == 5;
    P.riemann = d->riemann;
    P.tab     = c->d_tab;
    P.gas     = gas;
    P.bc      = bc;
    P.curved  = d->geometry != 0;
    const GeomP g = geom_of(*d);
    {
        // affine constants exactly as the reference composes them
        const VolM m = composite_metrics(g, d->nx, d->ny, 0.0, 0.0);
        P.xi_x  = m.xi_x;
        P.eta_y = m.eta_y;
        P.sc_x  = std::sqrt(m.xi_x * m.xi_x + m.xi_y * m.xi_y);
        P.sc_y  = std::sqrt(m.eta_x * m.eta_x + m.eta_y * m.eta_y);
    }
    if (P.curved)
    {
        const long long nv  = c->E * n * n;
        const long long nfx = static_cast<long long>(d->ny) * (d->nx + 1) * n;
        const long long nfy = static_cast<long long>(d->ny + 1) * d->nx * n;
        if (hipMalloc(&c->d_vmet, sizeof(double) * 4 * nv) != hipSuccess) return fail(FRSDG_ERR_HIP);
        if (hipMalloc(&c->d_fxm, sizeof(double) * 4 * nfx) != hipSuccess) return fail(FRSDG_ERR_HIP);
        if (hipMalloc(&c->d_fym, sizeof(double) * 4 * nfy) != hipSuccess) return fail(FRSDG_ERR_HIP);
        hipLaunchKernelGGL(k_geom_vol, dim3(static_cast<unsigned>((nv + 255) / 256)), dim3(256), 0, nullptr, g,
                           d->nx, d->ny, n, c->d_csi, c->E, c->d_vmet);
        hipLaunchKernelGGL(k_geom_face, dim3(static_cast<unsigned>((nfx + 255) / 256)), dim3(256), 0, nullptr, g,
                           d->nx, d->ny, n, c->d_csi, 1, c->d_fxm);
        hipLaunchKernelGGL(k_geom_face, dim3(static_cast<unsigned>((nfy + 255) / 256)), dim3(256), 0, nullptr, g,
                           d->nx, d->ny, n, c->d_csi, 0, c->d_fym);
        if (hipDeviceSynchronize() != hipSuccess) return fail(FRSDG_ERR_HIP);
        P.vmet = c->d_vmet;
        P.fxm  = c->d_fxm;
        P.fym  = c->d_fym;
    }
    if (d->coord_sys == 1)
    {
        const long long tot = c->E * n * n;
        if (hipMalloc(&c->d_rinv, sizeof(double) * tot) != hipSuccess) return fail(FRSDG_ERR_HIP);
        hipLaunchKernelGGL(k_rinv2d, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, nullptr, g, d->nx,
                           d->ny, n, c->d_csi, c->E, c->d_rinv);
        if (hipDeviceSynchronize() != hipSuccess) return fail(FRSDG_ERR_HIP);
        P.rinv = c->d_rinv;
    }
    (void) rc;
    *out = c;
    return FRSDG_OK;
}

int frsdg_destroy(frsdg_ctx* c)
{
    if (!c) return FRSDG_OK;
    free_solver(c);
    if (c->gexec) (void) hipGraphExecDestroy(c->gexec);
    if (c->graph) (void) hipGraphDestroy(c->graph);
    for (double* p : {c->d_tab, c->d_vmet, c->d_fxm, c->d_fym, c->d_csi, c->d_rinv, c->w1, c->w2, c->red, c->wb,
                      c->wbeta, c->adj_pool})
        if (p) (void) hipFree(p);
    for (int* p : {c->health, c->health_chk})
        if (p) (void) hipFree(p);
    delete c;
    return FRSDG_OK;
}

long long frsdg_ndofs(const frsdg_ctx* c) { return c ? c->nd : -1; }

int frsdg_get_tables(const frsdg_ctx* c, double* csi, double* w, double* D, double* dL, double* dR,
                     double* filt)
{
    if (!c) return FRSDG_ERR_INVALID;
    const int n = c->n;
    if (csi) std::memcpy(csi, c->tb.csi.data(), sizeof(double) * n);
    if (w) std::memcpy(w, c->tb.w.data(), sizeof(double) * n);
    if (D) std::memcpy(D, c->tb.D.data(), sizeof(double) * n * n);
    if (dL) std::memcpy(dL, c->tb.dL.data(), sizeof(double) * n);
    if (dR) std::memcpy(dR, c->tb.dR.data(), sizeof(double) * n);
    if (filt) std::memcpy(filt, c->tb.filt.data(), sizeof(double) * n * n);
    return FRSDG_OK;
}

int frsdg_residual(frsdg_ctx* c, double, const double* u, double* R, void* stream)
{
    if (!c || !u || !R) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(R)) return FRSDG_ERR_ALIGN;
    IORes io{u, R, nullptr, 1.0};
    return launch_res(c, io, static_cast<hipStream_t>(stream));
}

int frsdg_rhs(frsdg_ctx* c, double, const double* u, double* rhs, void* stream)
{
    if (!c || !u || !rhs) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(rhs)) return FRSDG_ERR_ALIGN;
    IORes io{u, rhs, nullptr, -1.0};
    return launch_res(c, io, static_cast<hipStream_t>(stream));
}

int frsdg_ifunction(frsdg_ctx* c, double, const double* u, const double* udot, double* F, void* stream)
{
    if (!c || !u || !udot || !F) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(udot) || !aligned(F)) return FRSDG_ERR_ALIGN;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c->d.cond_sys)
    {
        // fun_implicit with the change of variables: F = U_t + R(C(W))
        int rc = ensure_work(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p2c, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, c->w1);
        FR_TRY(hipGetLastError());
        u = c->w1;
    }
    IORes io{u, F, udot, 1.0};
    return launch_res(c, io, s);
}

int frsdg_jvp(frsdg_ctx* c, double, const double* u, const double* v, double sigma, double* out,
              void* stream)
{
    if (!c || !u || !v || !out) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(v) || !aligned(out)) return FRSDG_ERR_ALIGN;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c->d.cond_sys)
    {
        // (sigma I + J_R(C(W))) dC/dW v: the map's tangent, then the residual's
        int rc = ensure_work(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p2c_tan, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, v, c->w1, c->w2);
        FR_TRY(hipGetLastError());
        u = c->w1;
        v = c->w2;
    }
    IOTan io{u, v, out, sigma};
    if (c->d.dims == 2) return by_n<Tan2>(c->n, c, io, s);
    return by_n<Launch<IOTan>::template R1>(c->n, c, io, s);
}

int frsdg_jvp_cstep(frsdg_ctx* c, double, const double* u, const double* v, double sigma, double h,
                    double* out, void* stream)
{
    if (!c || !u || !v || !out || !(h > 0.0)) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(v) || !aligned(out)) return FRSDG_ERR_ALIGN;
    // the complex residual sees lambda-bound BCs frozen (boundary_conditions.hpp:270-294)
    if (c->d.cond_sys)
    {
        // Q = C(W + i h V) in complex arithmetic (petsc_solver_impl.hpp:297-326),
        // handed over as (Re Q, Im Q / h)
        int rc = ensure_work(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p2c_cplx, dim3(node_grid(c)), dim3(256), 0, static_cast<hipStream_t>(stream), c->nm, u,
                           v, h, 1.0 / h, c->w1, c->w2);
        FR_TRY(hipGetLastError());
        u = c->w1;
        v = c->w2;
    }
    const int keep      = c->p2.bc.freeze;
    c->p2.bc.freeze     = 1;
    c->p1.bc.freeze     = 1;
    IOCs io{u, v, out, sigma, h, 1.0 / h};
    const int rc    = launch_res(c, io, static_cast<hipStream_t>(stream));
    c->p2.bc.freeze = keep;
    c->p1.bc.freeze = keep;
    return rc;
}

int frsdg_vjp(frsdg_ctx* c, double, const double* u, const double* w, double sigma, double* out,
              void* stream)
{
    if (!c || !u || !w || !out) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(w) || !aligned(out)) return FRSDG_ERR_ALIGN;
    if (out == u || out == w) return FRSDG_ERR_INVALID; // neighbours' entries are still being read
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c->d.dims == 2 && !c->wb) FR_TRY(hipMalloc(&c->wb, sizeof(double) * 16 * c->n * c->E));
    if (c->d.cond_sys)
    {
        // (dC/dW)^T (sigma I + J_R(C(W))^T) w
        int rc = ensure_work(c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p2c, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, c->w1);
        FR_TRY(hipGetLastError());
        const AdjOut ao{sigma, 1.0, 0.0, 0.0, nullptr, nullptr};
        rc = c->d.dims == 1 ? by_n<Adj1>(c->n, c, c->w1, w, ao, c->w2, s)
                            : by_n<Adj2>(c->n, c, c->w1, w, ao, c->w2, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p2c_vjp, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, c->w2, out);
        return hip_rc(hipGetLastError());
    }
    const AdjOut ao{sigma, 1.0, 0.0, 0.0, nullptr, nullptr};
    if (c->d.dims == 1) return by_n<Adj1>(c->n, c, u, w, ao, out, s);
    return by_n<Adj2>(c->n, c, u, w, ao, out, s);
}

int frsdg_pre_jv_eval(frsdg_ctx* c, double t, const double* u, double sigma)
{
    if (!c || !u) return FRSDG_ERR_INVALID;
    c->jv_u     = u;
    c->jv_sigma = sigma;
    c->jv_t     = t;
    return FRSDG_OK;
}

int frsdg_matmult(frsdg_ctx* c, const double* v, double* out, void* stream)
{
    if (!c || !c->jv_u) return FRSDG_ERR_INVALID;
    if (c->d.jac_type == 1) return frsdg_jvp_cstep(c, c->jv_t, c->jv_u, v, c->jv_sigma, 1e-7, out, stream);
    return frsdg_jvp(c, c->jv_t, c->jv_u, v, c->jv_sigma, out, stream);
}

int frsdg_filter(frsdg_ctx* c, const double* u, double* out, void* stream)
{
    if (!c || !u || !out) return FRSDG_ERR_INVALID;
    if (!aligned(u) || !aligned(out)) return FRSDG_ERR_ALIGN;
    return do_filter(c, u, out, static_cast<hipStream_t>(stream));
}

int frsdg_filter_transpose(frsdg_ctx* c, const double* w, double* out, void* stream)
{
    if (!c || !w || !out) return FRSDG_ERR_INVALID;
    if (!aligned(w) || !aligned(out)) return FRSDG_ERR_ALIGN;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (weno_only(c)) return FRSDG_ERR_UNSUPPORTED; // the limiter's branches have no transpose here
    if (!c->d.apply_filter)
    {
        if (out != w) FR_TRY(hipMemcpyAsync(out, w, sizeof(double) * c->nd, hipMemcpyDeviceToDevice, s));
        return FRSDG_OK;
    }
    return by_n<Filt>(c->n, c, w, out, s, 1);
}

int frsdg_mass(frsdg_ctx* c, double* m, void* stream)
{
    if (!c || !m) return FRSDG_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n = c->n;
    if (c->d.dims == 2)
        hipLaunchKernelGGL(k_mass2d, dim3(static_cast<unsigned>((c->E * n * n + 255) / 256)), dim3(256), 0, s,
                           geom_of(c->d), c->d.nx, c->d.ny, n, c->d_csi, c->d_csi + n, c->E, m);
    else
        hipLaunchKernelGGL(k_mass1d, dim3(static_cast<unsigned>((c->nd + 255) / 256)), dim3(256), 0, s, c->d.nx, n,
                           c->d.lx / c->d.nx, c->d_csi + n, m);
    return hip_rc(hipGetLastError());
}

int frsdg_boundary_weights(frsdg_ctx* c, int dir, double* out, void* stream)
{
    if (!c || !out || c->d.dims != 2 || dir < 0 || dir > 3) return FRSDG_ERR_INVALID;
    const int along = (dir < 2) ? c->d.ny : c->d.nx;
    const long long tot = static_cast<long long>(along) * c->n;
    hipLaunchKernelGGL(k_wall2d, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), geom_of(c->d), c->d.nx, c->d.ny, c->n, c->d_csi,
                       c->d_csi + c->n, dir, out);
    return hip_rc(hipGetLastError());
}

int frsdg_apply_stabilization(frsdg_ctx* c, double* u, void* stream)
{
    if (!c || !u) return FRSDG_ERR_INVALID;
    if (!c->d.apply_filter && !c->d.apply_weno) return FRSDG_OK; // filters_active() (filter.hpp:710)
    return frsdg_filter(c, u, u, stream);
}

int frsdg_ssprk_step(frsdg_ctx* c, double, double dt, double* u, int stages, void* stream)
{
    if (!c || !u) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    if (stages != 3 && stages != 4) return FRSDG_ERR_INVALID;
    int rc = ensure_work(c);
    if (rc) return rc;
    return do_step(c, dt, u, stages, static_cast<hipStream_t>(stream));
}

int frsdg_ssprk_run(frsdg_ctx* c, double t, double dt, double* u, int stages, int nsteps, void* stream)
{
    if (!c || !u || nsteps < 0) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    if (stages != 3 && stages != 4) return FRSDG_ERR_INVALID;
    int rc = ensure_work(c);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    constexpr int CH = 16; // steps per captured graph
    // the health flags of this run (outside any capture, so a memset is fine)
    FR_TRY(hipMemsetAsync(c->health, 0, sizeof(int), s));
    auto healthy = [&]() -> int
    {
        int f = 0;
        FR_TRY(hipMemcpyAsync(&f, c->health, sizeof(int), hipMemcpyDeviceToHost, s));
        FR_TRY(hipStreamSynchronize(s));
        return (f & c->stop_mask) ? FRSDG_ERR_NUMERIC : FRSDG_OK;
    };
    if (nsteps < CH) // short runs: plain launches
    {
        for (int k = 0; k < nsteps; ++k)
            if ((rc = frsdg_ssprk_step(c, t + k * dt, dt, u, stages, stream))) return rc;
        return healthy();
    }
    if (!c->gexec || c->g_u != u || c->g_dt != dt || c->g_st != stages)
    {
        if (c->gexec) (void) hipGraphExecDestroy(c->gexec);
        if (c->graph) (void) hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
        hipStream_t cs;
        FR_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
        {
            (void) hipStreamDestroy(cs);
            return FRSDG_ERR_HIP;
        }
        int rr = FRSDG_OK;
        for (int k = 0; k < CH && !rr; ++k) rr = do_step(c, dt, u, stages, cs);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        (void) hipStreamDestroy(cs);
        if (rr || ec != hipSuccess) return rr ? rr : FRSDG_ERR_HIP;
        c->graph = g;
        FR_TRY(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->g_u  = u;
        c->g_dt = dt;
        c->g_st = stages;
    }
    const int full = nsteps / CH, rem = nsteps % CH;
    // -ts_error_if_step_fails: stop after the first graph of steps whose stage
    // outputs raised a flag of the stop mask (checked once per graph)
    for (int k = 0; k < full; ++k)
    {
        FR_TRY(hipGraphLaunch(c->gexec, s));
        if ((rc = healthy())) return rc;
    }
    for (int k = 0; k < rem; ++k)
        if ((rc = do_step(c, dt, u, stages, s))) return rc;
    return healthy();
}

int frsdg_set_stop_mask(frsdg_ctx* c, int mask)
{
    if (!c || (mask & ~7)) return FRSDG_ERR_INVALID;
    c->stop_mask = mask;
    return FRSDG_OK;
}

int frsdg_step_health(frsdg_ctx* c, int* flags, int clear, void* stream)
{
    if (!c || !flags) return FRSDG_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    *flags = 0;
    if (!c->health) return FRSDG_OK;
    FR_TRY(hipMemcpyAsync(flags, c->health, sizeof(int), hipMemcpyDeviceToHost, s));
    if (clear) FR_TRY(hipMemsetAsync(c->health, 0, sizeof(int), s));
    FR_TRY(hipStreamSynchronize(s));
    return FRSDG_OK;
}

int frsdg_check_state(frsdg_ctx* c, const double* u, int* flags, void* stream)
{
    if (!c || !u || !flags) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!c->health_chk) FR_TRY(hipMalloc(&c->health_chk, sizeof(int)));
    FR_TRY(hipMemsetAsync(c->health_chk, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_health, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, c->health_chk);
    FR_TRY(hipGetLastError());
    FR_TRY(hipMemcpyAsync(flags, c->health_chk, sizeof(int), hipMemcpyDeviceToHost, s));
    FR_TRY(hipStreamSynchronize(s));
    return FRSDG_OK;
}

int frsdg_impose_ic(frsdg_ctx* c, int kind, const double* prm, double* u, void* stream)
{
    if (!c || !prm || !u || kind < 0 || kind > 4) return FRSDG_ERR_INVALID;
    ICP ic{};
    ic.kind = kind;
    for (int q = 0; q < 8; ++q) ic.prm[q] = prm[q];
    ic.t = 0.0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n = c->n;
    if (c->d.dims == 1)
    {
        if (kind != 0 && kind != 3) return FRSDG_ERR_INVALID;
        const long long tot = static_cast<long long>(c->d.nx) * n;
        hipLaunchKernelGGL(k_ic1d, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, s, c->p1.gas, ic,
                           c->d.nx, n, c->d.lx, c->d_csi, u);
        if (c->d.cond_sys) hipLaunchKernelGGL(k_c2p, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, u);
        return hip_rc(hipGetLastError());
    }
    if (kind == 3) return FRSDG_ERR_INVALID;
    const long long tot = c->E * n * n;
    hipLaunchKernelGGL(k_ic2d, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, s, geom_of(c->d),
                       c->p2.gas, ic, c->d.geometry == 0 ? 1 : 0, c->d.nx, c->d.ny, n, c->d.lx, c->d.ly,
                       c->d_csi, c->E, u);
    // impose_ic stores cons_to_prim of the state with cond_sys (physics_impl_2D.hpp:180-190)
    if (c->d.cond_sys) hipLaunchKernelGGL(k_c2p, dim3(node_grid(c)), dim3(256), 0, s, c->nm, u, u);
    return hip_rc(hipGetLastError());
}

int frsdg_l2_error(frsdg_ctx* c, int kind, const double* prm, double t, const double* u, double* out,
                   void* stream)
{
    if (!c || !prm || !u || !out || c->d.dims != 2 || kind < 0 || kind > 4 || kind == 3) return FRSDG_ERR_INVALID;
    ICP ic{};
    ic.kind = kind;
    for (int q = 0; q < 8; ++q) ic.prm[q] = prm[q];
    ic.t = t;
    hipStream_t     s   = static_cast<hipStream_t>(stream);
    const int       n   = c->n;
    const long long tot = c->E * n * n;
    const long long nb  = (tot + 255) / 256;
    if (!c->red) FR_TRY(hipMalloc(&c->red, sizeof(double) * 4 * (nb + 1)));
    return l2_error_into(c, ic, u, c->red, out, s);
}

int frsdg_l2_error_noalloc(frsdg_ctx* c, int kind, const double* prm, double t, const double* u, double* out,
                           void* stream)
{
    if (!c || !prm || !u || !out || c->d.dims != 2 || kind < 0 || kind > 4 || kind == 3) return FRSDG_ERR_INVALID;
    const long long nb = (c->E * c->n * c->n + 255) / 256;
    if (nb > FRSDG_L2_SCRATCH_BLOCKS) return FRSDG_ERR_INVALID; // beyond the static scratch (E n^2 > 2^24 nodes)
    double* scr = nullptr;
    int     rc  = scratch_device_ptr(&scr);
    if (rc) return rc;
    ICP ic{};
    ic.kind = kind;
    for (int q = 0; q < 8; ++q) ic.prm[q] = prm[q];
    ic.t = t;
    return l2_error_into(c, ic, u, scr, out, static_cast<hipStream_t>(stream));
}

} // extern "C"

// ---------------- unsteady discrete adjoint (binomial checkpointing) ----------------
namespace
{
    // Smallest r with C(s + r, s) >= T (repetitions of the binomial schedule).
    int revolve_repetitions(long long T, int s)
    {
        int r = 0;
        for (;; ++r)
        {
            // C(s + r, s) with early exit once it reaches T
            double c = 1.0;
            for (int i = 1; i <= r; ++i) c = c * (s + i) / i;
            if (c >= static_cast<double>(T) - 0.5) return r;
        }
    }
    long long binom(int n, int k)
    {
        if (k < 0 || k > n) return 0;
        double c = 1.0;
        for (int i = 1; i <= k; ++i) c = c * (n - k + i) / i;
        return static_cast<long long>(c + 0.5);
    }

    // One reverse sweep of frsdg_unsteady_adjoint: the pool holds `slots`
    // snapshot states, the recompute work state, the stage states w1/w2 of the
    // step being reversed and the adjoint stage vectors l3, l2, l1.
    struct UnsteadyAdj
    {
        frsdg_ctx*          c;
        hipStream_t         st;
        double              dt;
        double*             pool;
        int                 slots;
        long long           nd;
        frsdg_adjoint_stats* S;
        int                 live = 0;

        double* slot(int i) const { return pool + static_cast<long long>(i) * nd; }
        double* work() const { return slot(slots); }
        double* w1() const { return slot(slots + 1); }
        double* w2() const { return slot(slots + 2); }
        double* l3() const { return slot(slots + 3); }
        double* l2() const { return slot(slots + 4); }
        double* l1() const { return slot(slots + 5); }

        int copy(double* dst, const double* src)
        {
            return hip_rc(hipMemcpyAsync(dst, src, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
        }
        int advance(double* x, long long k)
        {
            for (long long i = 0; i < k; ++i)
            {
                int rc = do_step(c, dt, x, 3, st);
                if (rc) return rc;
            }
            S->forward_steps += static_cast<int>(k);
            return FRSDG_OK;
        }
        int adj(const double* u, const double* w, const AdjOut& ao, double* out)
        {
            if (c->d.dims == 2) return by_n<Adj2>(c->n, c, u, w, ao, out, st);
            return by_n<Adj1>(c->n, c, u, w, ao, out, st);
        }
        // lam <- (d u_{k+1} / d u_k)^T lam for the step that starts at u (SSPRK(3,3)
        // with rhs = -R, then the post-step filter F):
        //   u1 = u - dt R(u), u2 = 3/4 u + 1/4 (u1 - dt R(u1)), u+ = F (1/3 u + 2/3 (u2 - dt R(u2)));
        //   l3 = F^T lam, l2 = 2/3 l3 - 2/3 dt J(u2)^T l3, l1 = 1/4 l2 - 1/4 dt J(u1)^T l2,
        //   lam = l1 - dt J(u)^T l1 + 1/3 l3 + 3/4 l2.
        int step_back(const double* u, double* lam)
        {
            int rc;
            {
                IOStage a{u, u, w1(), 0.0, 1.0, dt};
                if ((rc = by_n<Stage>(c->n, c, a, 0, st))) return rc;
                IOStage b{w1(), u, w2(), 0.75, 0.25, dt};
                if ((rc = by_n<Stage>(c->n, c, b, 0, st))) return rc;
            }
            // l3 = F^T lam (a copy of lam without the filter: lam is overwritten last)
            if (c->d.apply_filter)
                rc = by_n<Filt>(c->n, c, lam, l3(), st, 1);
            else
                rc = copy(l3(), lam);
            if (rc) return rc;
            if ((rc = adj(w2(), l3(), AdjOut{2.0 / 3.0, -2.0 / 3.0 * dt, 0.0, 0.0, nullptr, nullptr}, l2()))) return rc;
            if ((rc = adj(w1(), l2(), AdjOut{0.25, -0.25 * dt, 0.0, 0.0, nullptr, nullptr}, l1()))) return rc;
            if ((rc = adj(u, l1(), AdjOut{1.0, -dt, 1.0 / 3.0, 0.75, l3(), l2()}, lam))) return rc;
            S->adjoint_steps++;
            return FRSDG_OK;
        }
        // reverse steps a..b-1 starting from the state ua at step a (never modified)
        int reverse(long long a, long long b, const double* ua, double* lam, int free)
        {
            const long long T = b - a;
            int rc;
            if (T == 1) return step_back(ua, lam);
            if (free == 0)
            {
                for (long long k = b - 1; k >= a; --k)
                {
                    if ((rc = copy(work(), ua)) || (rc = advance(work(), k - a)) || (rc = step_back(work(), lam)))
                        return rc;
                }
                return FRSDG_OK;
            }
            const int       r      = revolve_repetitions(T, free);
            const long long second = std::min(T - 1, binom(free - 1 + r, free - 1));
            const long long m      = b - second;
            double*         x      = slot(free - 1);
            if ((rc = copy(x, ua)) || (rc = advance(x, m - a))) return rc;
            S->snapshots_used = std::max(S->snapshots_used, ++live);
            rc = reverse(m, b, x, lam, free - 1);
            --live;
            if (rc) return rc;
            return reverse(a, m, ua, lam, free);
        }
    };
} // namespace

extern "C" {

int frsdg_unsteady_adjoint(frsdg_ctx* c, double, double dt, int nsteps, int max_snapshots, const double* u0,
                           const double* lam_T, double* lam_0, frsdg_adjoint_stats* stats, void* stream)
{
    if (!c || !u0 || !lam_T || !lam_0 || nsteps < 0 || max_snapshots < 0 || !(dt > 0.0)) return FRSDG_ERR_INVALID;
    if (!aligned(u0) || !aligned(lam_T) || !aligned(lam_0)) return FRSDG_ERR_ALIGN;
    if (lam_0 == u0) return FRSDG_ERR_INVALID;
    // the WENO limiter (apply_weno without the filter) has no transpose
    if (c->d.apply_weno && !c->d.apply_filter) return FRSDG_ERR_UNSUPPORTED;
    hipStream_t st = static_cast<hipStream_t>(stream);
    frsdg_adjoint_stats local{};
    frsdg_adjoint_stats* S = stats ? stats : &local;
    *S = frsdg_adjoint_stats{};
    FR_TRY(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    int rc = ensure_work(c);
    if (rc) return rc;
    if (c->d.dims == 2 && !c->wb) FR_TRY(hipMalloc(&c->wb, sizeof(double) * 16 * c->n * c->E));
    const int slots = static_cast<int>(std::min<long long>(max_snapshots, std::max(nsteps - 1, 0)));
    const long long need = static_cast<long long>(slots + 6);
    if (c->adj_pool_states < need)
    {
        if (c->adj_pool) FR_TRY(hipFree(c->adj_pool));
        c->adj_pool        = nullptr;
        c->adj_pool_states = 0;
        FR_TRY(hipMalloc(&c->adj_pool, sizeof(double) * need * c->nd));
        c->adj_pool_states = need;
    }
    if (lam_0 != lam_T) FR_TRY(hipMemcpyAsync(lam_0, lam_T, sizeof(double) * c->nd, hipMemcpyDeviceToDevice, st));
    if (nsteps > 0)
    {
        UnsteadyAdj A{c, st, dt, c->adj_pool, slots, c->nd, S};
        if ((rc = A.reverse(0, nsteps, u0, lam_0, slots))) return rc;
    }
    FR_TRY(hipStreamSynchronize(st));
    S->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return FRSDG_OK;
}

} // extern "C"

// fr_oracle.cpp — CPU ORACLE of the FR residual path (test infrastructure only).
// See fr_oracle.hpp for what this is and is not.  The solver classes below
// restate physics_impl_2D.hpp / physics_impl_1D.hpp / physics_evaluator.hpp /
// geometry_models.hpp / boundary_conditions.hpp / petsc_solver_impl.hpp of the
// reference for a single rank (periodic neighbours are index wraps instead of
// DMDA ghosts).
#include "fr_oracle.h"
#include "fr_oracle.hpp"

#include <omp.h>

#include <cstring>
#include <memory>
#include <numbers>

namespace fro
{
    // ------------------------------------------------------------------
    // geometry_models.hpp:19-67 containers
    // ------------------------------------------------------------------
    struct GeomJac
    {
        double dx_dcsi{1.0}, dy_dcsi{0.0}, dx_deta{0.0}, dy_deta{1.0};
    };
    struct VolMetrics
    {
        double xi_x{1.0}, xi_y{0.0}, eta_x{0.0}, eta_y{1.0}, detJ{1.0};
        double norm_grad_xi() const { return std::sqrt(xi_x * xi_x + xi_y * xi_y); }
        double norm_grad_eta() const { return std::sqrt(eta_x * eta_x + eta_y * eta_y); }
    };
    struct FaceN
    {
        double nx{1.0}, ny{0.0};
    };

    // One struct, one switch: the reference selects the policy at compile time
    // (ProblemTraits::GeomModel); the restatement selects it at run time.
    struct Geometry
    {
        int    kind{0};
        double Lx{1.0}, Ly{1.0};
        double prm[4]{};

        // tests/bump.cpp:29-119 wedge profile (height, slope), Lx = m_L_x
        std::pair<double, double> wedge(double x) const
        {
            const double x_start      = 0.15 * Lx;
            const double x_c_straight = 0.30 * Lx;
            const double x_e_start    = 0.55 * Lx;
            const double x_e_end      = 0.70 * Lx;
            const double compress_slope = 0.2;
            const double expand_slope   = -0.2;
            const double A_c     = compress_slope / (2.0 * (x_c_straight - x_start));
            const double dx_c    = x_c_straight - x_start;
            const double h_c_end = A_c * dx_c * dx_c;
            if (x <= x_start) return {0.0, 0.0};
            if (x <= x_c_straight)
            {
                const double dx = x - x_start;
                return {-(A_c * dx * dx), -(2.0 * A_c * dx)};
            }
            if (x <= x_e_start)
            {
                const double dx = x - x_c_straight;
                return {-(h_c_end + compress_slope * dx), -compress_slope};
            }
            const double dx_run    = x_e_start - x_c_straight;
            const double h_e_start = h_c_end + compress_slope * dx_run;
            const double dx_e      = x_e_end - x_e_start;
            const double A_e       = (expand_slope - compress_slope) / (2.0 * dx_e);
            if (x <= x_e_end)
            {
                const double dx = x - x_e_start;
                return {-(h_e_start + compress_slope * dx + A_e * dx * dx),
                        -(compress_slope + 2.0 * A_e * dx)};
            }
            const double h_e_end = h_e_start + compress_slope * dx_e + A_e * dx_e * dx_e;
            const double dx      = x - x_e_end;
            return {-(h_e_end + expand_slope * dx), -expand_slope};
        }
        // Channel bump: y_w = A sin^2(pi (x - xc + w/2)) for |x - xc| <= w/2.
        // prm = {A, x0 (left end of domain), xc, w}
        std::pair<double, double> bump(double x) const
        {
            const double A = prm[0], xc = prm[2], w = prm[3];
            if (std::fabs(x - xc) > 0.5 * w) return {0.0, 0.0};
            const double s = std::numbers::pi * (x - xc + 0.5 * w) / w;
            return {A * std::sin(s) * std::sin(s),
                    A * 2.0 * std::sin(s) * std::cos(s) * std::numbers::pi / w};
        }

        std::pair<double, double> phys(double csi, double eta) const
        {
            switch (kind)
            {
                case 0: return {Lx * csi, Ly * eta}; // geometry_models.hpp:100-103
                case 1:
                {
                    // geometry_models.hpp:162-173 (straight profile h=0.5)
                    const double x_phys = Lx * csi;
                    const double h_x    = 0.5;
                    return {x_phys, -h_x * (1.0 - eta) + h_x * eta};
                }
                case 2:
                {
                    // tests/bump.cpp:121-131
                    const double x_phys = Lx * csi;
                    const auto [h_x, s] = wedge(x_phys);
                    (void) s;
                    return {x_phys, h_x * (1.0 - eta) + Ly * eta};
                }
                case 3:
                {
                    const double x_phys = prm[1] + Lx * csi;
                    const auto [h_x, s] = bump(x_phys);
                    (void) s;
                    return {x_phys, h_x * (1.0 - eta) + Ly * eta};
                }
                default:
                {
                    const double A = prm[0];
                    const double X = prm[1] + Lx * csi, Y = prm[2] + Ly * eta;
                    const double s = std::sin(std::numbers::pi * X) * std::sin(std::numbers::pi * Y);
                    return {X + A * s, Y + A * s};
                }
            }
        }
        GeomJac jac(double csi, double eta) const
        {
            switch (kind)
            {
                case 0: return {Lx, 0.0, 0.0, Ly}; // :305-311
                case 1:
                {
                    // :375-390 — metric assumes y = h*eta (quirk Q4, kept)
                    const double h_x = 0.5, dh_dx = 0.0;
                    return {Lx, dh_dx * Lx * eta, 0.0, h_x};
                }
                case 2:
                {
                    // tests/bump.cpp:133-146
                    const auto [h_x, dh_dx] = wedge(Lx * csi);
                    return {Lx, dh_dx * Lx * (1.0 - eta), 0.0, Ly - h_x};
                }
                case 3:
                {
                    const auto [h_x, dh_dx] = bump(prm[1] + Lx * csi);
                    return {Lx, dh_dx * Lx * (1.0 - eta), 0.0, Ly - h_x};
                }
                default:
                {
                    const double A = prm[0], pi = std::numbers::pi;
                    const double X = prm[1] + Lx * csi, Y = prm[2] + Ly * eta;
                    const double cx = std::cos(pi * X), sx = std::sin(pi * X);
                    const double cy = std::cos(pi * Y), sy = std::sin(pi * Y);
                    const double ds_dX = pi * cx * sy, ds_dY = pi * sx * cy;
                    return {Lx * (1.0 + A * ds_dX), Lx * A * ds_dX, Ly * A * ds_dY,
                            Ly * (1.0 + A * ds_dY)};
                }
            }
        }
        FaceN x_face(double csi, double eta) const
        {
            if (kind == 4)
            {
                // warped: n = grad(xi)/|grad(xi)| ~ (y_eta, -x_eta)
                const auto   J  = jac(csi, eta);
                const double in = 1.0 / std::sqrt(J.dy_deta * J.dy_deta + J.dx_deta * J.dx_deta);
                return {J.dy_deta * in, -J.dx_deta * in};
            }
            return {1.0, 0.0}; // :313-316, :392-395, bump.cpp:148-153
        }
        FaceN y_face(double csi, double eta) const
        {
            switch (kind)
            {
                case 0: return {0.0, 1.0}; // :317-320
                case 1:
                {
                    // :397-408
                    const double dh_dx = 0.0;
                    double       unx = -dh_dx * eta, uny = 1.0;
                    double       inv = 1.0 / std::sqrt(unx * unx + uny * uny);
                    return {unx * inv, uny * inv};
                }
                case 2:
                case 3:
                {
                    // tests/bump.cpp:155-170
                    const double x_phys = (kind == 2 ? 0.0 : prm[1]) + Lx * csi;
                    const double dh_dx  = (kind == 2 ? wedge(x_phys) : bump(x_phys)).second;
                    double       unx    = -dh_dx * Lx * (1.0 - eta);
                    double       uny    = Lx;
                    double       inv    = 1.0 / std::sqrt(unx * unx + uny * uny);
                    return {unx * inv, uny * inv};
                }
                default:
                {
                    const auto   J  = jac(csi, eta);
                    const double in = 1.0 / std::sqrt(J.dy_dcsi * J.dy_dcsi + J.dx_dcsi * J.dx_dcsi);
                    return {-J.dy_dcsi * in, J.dx_dcsi * in};
                }
            }
        }
    };

    // boundary_conditions.hpp:8-23
    enum Dir
    {
        LEFT = 0,
        RIGHT,
        TOP,
        BOTTOM
    };
    enum BCKind
    {
        BC_WALL = 0,
        BC_OUTFLOW = 1,
        BC_DIRICHLET = 2,
        BC_PERIODIC = 5,
        BC_SLIPWALL = 6
    };

    template <class U>
    using St4 = std::array<U, 4>;

    // boundary_conditions.hpp:214-324 — per-face std::function maps.  Real
    // and complex variants; lambda-bound faces (Dirichlet, curved slip wall)
    // drop imaginary parts (bind_lambda :270-294).
    struct BCs2D
    {
        using RealFunc = std::function<St4<double>(double, double, double, FaceN, const St4<double>&)>;
        using CplxFunc = std::function<St4<cplx>(double, double, double, FaceN, const St4<cplx>&)>;
        int                     type[4]{1, 1, 1, 1};
        std::array<RealFunc, 4> real;
        std::array<CplxFunc, 4> cpx;

        void bind(int dir, int t, const double* fixed)
        {
            type[dir] = t;
            if (t == BC_WALL)
            {
                const std::size_t axis = dir / 2; // :300-308
                real[dir] = [axis](double, double, double, FaceN, const St4<double>& st)
                {
                    auto r       = st;
                    r[1 + axis] = -r[1 + axis];
                    return r;
                };
                cpx[dir] = [axis](double, double, double, FaceN, const St4<cplx>& st)
                {
                    auto r       = st;
                    r[1 + axis] = -r[1 + axis];
                    return r;
                };
                return;
            }
            if (t == BC_OUTFLOW || t == BC_PERIODIC)
            {
                real[dir] = [](double, double, double, FaceN, const St4<double>& st) { return st; };
                cpx[dir]  = [](double, double, double, FaceN, const St4<cplx>& st) { return st; };
                return;
            }
            RealFunc f;
            if (t == BC_DIRICHLET)
            {
                St4<double> s{fixed[0], fixed[1], fixed[2], fixed[3]};
                f = [s](double, double, double, FaceN, const St4<double>&) { return s; };
            }
            else
            {
                // tests/bump.cpp:250-272 curvilinear slip wall: reflect the
                // velocity about the face normal of the geometry model.
                f = [](double, double, double, FaceN n, const St4<double>& in)
                {
                    auto         ex    = in;
                    const double u_in  = in[1] / in[0];
                    const double v_in  = in[2] / in[0];
                    const double vn_in = u_in * n.nx + v_in * n.ny;
                    const double u_out = u_in - 2.0 * vn_in * n.nx;
                    const double v_out = v_in - 2.0 * vn_in * n.ny;
                    ex[1]              = ex[0] * u_out;
                    ex[2]              = ex[0] * v_out;
                    return ex;
                };
            }
            type[dir] = t;
            real[dir] = f;
            cpx[dir]  = [f](double tt, double x, double y, FaceN n, const St4<cplx>& c)
            {
                St4<double> rs{std::real(c[0]), std::real(c[1]), std::real(c[2]), std::real(c[3])};
                auto        r = f(tt, x, y, n, rs);
                return St4<cplx>{cplx{r[0]}, cplx{r[1]}, cplx{r[2]}, cplx{r[3]}};
            };
        }
        template <class U>
        St4<U> eval(int dir, double t, double x, double y, FaceN n, const St4<U>& in) const
        {
            if constexpr (is_cplx<U>::value)
                return cpx[dir](t, x, y, n, in);
            else
                return real[dir](t, x, y, n, in);
        }
    };

    // ------------------------------------------------------------------
    // 2D evaluator — physics_evaluator.hpp:158-318 + physics_impl_2D.hpp
    // ------------------------------------------------------------------
    struct Solver2D
    {
        fro_desc  d;
        GLLTables tb;
        Gas       gas;
        Geometry  geo;
        BCs2D     bcs;
        bool      per_x{false}, per_y{false};
        int       n{0};
        long      ndofs_elem{0};

        explicit Solver2D(const fro_desc& dd) : d(dd)
        {
            tb  = make_tables(d.p, d.c_vcjh, d.filt_cutoff, d.filt_strength, d.filt_alpha);
            n   = tb.n;
            gas = Gas{d.gas_kind, d.gamma, d.R_spec, d.vdw_a, d.vdw_b};
            geo.kind = d.geometry;
            geo.Lx   = d.lx;
            geo.Ly   = d.ly;
            for (int i = 0; i < 4; ++i) geo.prm[i] = d.geom_param[i];
            for (int k = 0; k < 4; ++k) bcs.bind(k, d.bc_type[k], d.bc_state[k]);
            per_x      = d.bc_type[LEFT] == BC_PERIODIC && d.bc_type[RIGHT] == BC_PERIODIC;
            per_y      = d.bc_type[TOP] == BC_PERIODIC && d.bc_type[BOTTOM] == BC_PERIODIC;
            ndofs_elem = 4L * n * n;
        }
        long ndofs() const { return ndofs_elem * d.nx * d.ny; }

        // physics_evaluator.hpp:235-246
        double comp_x(long i, int k) const
        {
            const double h_x = 1.0 / static_cast<double>(d.nx);
            return h_x * (static_cast<double>(i) + 0.5 * (1.0 + tb.csi[k]));
        }
        double comp_y(long j, int l) const
        {
            const double h_y = 1.0 / static_cast<double>(d.ny);
            return h_y * (static_cast<double>(j) + 0.5 * (1.0 + tb.csi[l]));
        }
        // physics_evaluator.hpp:282-305
        VolMetrics metrics(double cx, double cy) const
        {
            const auto   gj      = geo.jac(cx, cy);
            const double dcsi_dr = 0.5 / static_cast<double>(d.nx);
            const double deta_ds = 0.5 / static_cast<double>(d.ny);
            const double dx_dr   = gj.dx_dcsi * dcsi_dr;
            const double dy_dr   = gj.dy_dcsi * dcsi_dr;
            const double dx_ds   = gj.dx_deta * deta_ds;
            const double dy_ds   = gj.dy_deta * deta_ds;
            const double detJ    = dx_dr * dy_ds - dx_ds * dy_dr;
            return {dy_ds / detJ, -dx_ds / detJ, -dy_dr / detJ, dx_dr / detJ, detJ};
        }

        // element span with single-rank periodic wrap (DMDA ghosts)
        template <class U>
        U* elem(U* base, long i, long j) const
        {
            if (i < 0) i += d.nx;
            if (i >= d.nx) i -= d.nx;
            if (j < 0) j += d.ny;
            if (j >= d.ny) j -= d.ny;
            return base + (j * d.nx + i) * ndofs_elem;
        }
        int idx(int eq, int k, int l) const { return (l * n + k) * 4 + eq; } // :27-35

        template <class U>
        St4<U> riemann(const St4<U>& UL, const St4<U>& UR, double nx, double ny) const
        {
            return d.riemann == 0 ? rusanov2(UL, UR, gas, nx, ny) : hllc2(UL, UR, gas, nx, ny);
        }

        // physics_impl_2D.hpp:267-357 compute_volume_term (Cartesian branch)
        template <class U>
        void volume_term(const U* u_C, U* res, long i, long j) const
        {
            const long nd = ndofs_elem;
            std::fill(res, res + nd, U{0.0});
            std::vector<U> F(nd, U{0.0}); // per-element heap scratch, as :285-286
            std::vector<U> G(nd, U{0.0});
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k)
                {
                    St4<U> s{u_C[idx(0, k, l)], u_C[idx(1, k, l)], u_C[idx(2, k, l)],
                             u_C[idx(3, k, l)]};
                    St4<U> fx, fy;
                    U      pres;
                    euler_fluxes_2d_w_p(s, gas, fx, fy, pres);
                    for (int eq = 0; eq < 4; ++eq)
                    {
                        F[idx(eq, k, l)] = fx[eq];
                        G[idx(eq, k, l)] = fy[eq];
                    }
                    if (d.coord_sys == 1) // radial source :306-319 (r = y/2 at the mapped node)
                    {
                        const auto   pos    = geo.phys(comp_x(i, k), comp_y(j, l));
                        const double radius = 0.5 * pos.second;
                        const double inv_r  = (radius > 1e-8) ? 1.0 / radius : 0.0;
                        for (int eq = 0; eq < 4; ++eq) res[idx(eq, k, l)] += G[idx(eq, k, l)] * inv_r;
                        res[idx(2, k, l)] -= pres * inv_r;
                    }
                }
            const auto& D = tb.D;
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k)
                {
                    const auto m = metrics(comp_x(i, k), comp_y(j, l));
                    for (int eq = 0; eq < 4; ++eq)
                    {
                        U dF_dxi{0.0}, dG_dxi{0.0};
                        for (int mm = 0; mm < n; ++mm)
                        {
                            dG_dxi += D[k * n + mm] * G[idx(eq, mm, l)];
                            dF_dxi += D[k * n + mm] * F[idx(eq, mm, l)];
                        }
                        U dG_deta{0.0}, dF_deta{0.0};
                        for (int mm = 0; mm < n; ++mm)
                        {
                            dG_deta += D[l * n + mm] * G[idx(eq, k, mm)];
                            dF_deta += D[l * n + mm] * F[idx(eq, k, mm)];
                        }
                        const U px = m.xi_x * dF_dxi + m.eta_x * dF_deta;
                        const U py = m.xi_y * dG_dxi + m.eta_y * dG_deta;
                        res[idx(eq, k, l)] += px + py;
                    }
                }
        }

        // physics_impl_2D.hpp:689-770 eval_x_edge
        template <class U>
        void x_edge(const U* x, U* f, long I, long j, bool updL, bool updR, bool bcL, bool bcR,
                    double t) const
        {
            const U*     u_L    = elem(x, I - 1, j);
            const U*     u_R    = elem(x, I, j);
            const double cx     = comp_x(I, 0);
            const auto&  dL     = tb.dL;
            const auto&  dR     = tb.dR;
            for (int l = 0; l < n; ++l)
            {
                const double cy   = comp_y(j, l);
                const auto   fg   = geo.x_face(cx, cy);
                const auto   m    = metrics(cx, cy);
                const double nx   = fg.nx, ny = fg.ny;
                const double sc   = m.norm_grad_xi();
                St4<U>       UL, UR;
                if (!bcR) UR = {u_R[idx(0, 0, l)], u_R[idx(1, 0, l)], u_R[idx(2, 0, l)], u_R[idx(3, 0, l)]};
                if (!bcL)
                    UL = {u_L[idx(0, n - 1, l)], u_L[idx(1, n - 1, l)], u_L[idx(2, n - 1, l)],
                          u_L[idx(3, n - 1, l)]};
                if (bcL)
                {
                    const auto [xp, yp] = geo.phys(cx, cy);
                    UL = bcs.eval(LEFT, t, xp, yp, fg, UR);
                }
                if (bcR)
                {
                    const auto [xp, yp] = geo.phys(cx, cy);
                    UR = bcs.eval(RIGHT, t, xp, yp, fg, UL);
                }
                St4<U> fxL, fyL, fxR, fyR;
                U      pL, pR;
                euler_fluxes_2d_w_p(UL, gas, fxL, fyL, pL);
                euler_fluxes_2d_w_p(UR, gas, fxR, fyR, pR);
                const auto Fs = riemann(UL, UR, nx, ny);
                for (int eq = 0; eq < 4; ++eq)
                {
                    if (updL)
                    {
                        U*      res_L = elem(f, I - 1, j);
                        const U nfL   = fxL[eq] * nx + fyL[eq] * ny;
                        for (int k = 0; k < n; ++k) res_L[idx(eq, k, l)] += sc * (Fs[eq] - nfL) * dR[k];
                    }
                    if (updR)
                    {
                        U*      res_R = elem(f, I, j);
                        const U nfR   = fxR[eq] * nx + fyR[eq] * ny;
                        for (int k = 0; k < n; ++k) res_R[idx(eq, k, l)] += sc * (Fs[eq] - nfR) * dL[k];
                    }
                }
            }
        }

        // physics_impl_2D.hpp:797-877 eval_y_edge
        template <class U>
        void y_edge(const U* x, U* f, long i, long J, bool updB, bool updT, bool bcB, bool bcT,
                    double t) const
        {
            const U*     u_B = elem(x, i, J - 1);
            const U*     u_T = elem(x, i, J);
            const double cy  = comp_y(J, 0);
            const auto&  dL  = tb.dL;
            const auto&  dR  = tb.dR;
            for (int k = 0; k < n; ++k)
            {
                const double cx = comp_x(i, k);
                const auto   fg = geo.y_face(cx, cy);
                const auto   m  = metrics(cx, cy);
                const double nx = fg.nx, ny = fg.ny;
                const double sc = m.norm_grad_eta();
                St4<U>       UB, UT;
                if (!bcT) UT = {u_T[idx(0, k, 0)], u_T[idx(1, k, 0)], u_T[idx(2, k, 0)], u_T[idx(3, k, 0)]};
                if (!bcB)
                    UB = {u_B[idx(0, k, n - 1)], u_B[idx(1, k, n - 1)], u_B[idx(2, k, n - 1)],
                          u_B[idx(3, k, n - 1)]};
                if (bcB)
                {
                    const auto [xp, yp] = geo.phys(cx, cy);
                    UB = bcs.eval(BOTTOM, t, xp, yp, fg, UT);
                }
                if (bcT)
                {
                    const auto [xp, yp] = geo.phys(cx, cy);
                    UT = bcs.eval(TOP, t, xp, yp, fg, UB);
                }
                St4<U> fxB, fyB, fxT, fyT;
                U      pB, pT;
                euler_fluxes_2d_w_p(UB, gas, fxB, fyB, pB);
                euler_fluxes_2d_w_p(UT, gas, fxT, fyT, pT);
                const auto Gs = riemann(UB, UT, nx, ny);
                for (int eq = 0; eq < 4; ++eq)
                {
                    if (updB)
                    {
                        U*      res_B = elem(f, i, J - 1);
                        const U nfB   = fxB[eq] * nx + fyB[eq] * ny;
                        for (int l = 0; l < n; ++l) res_B[idx(eq, k, l)] += sc * (Gs[eq] - nfB) * dR[l];
                    }
                    if (updT)
                    {
                        U*      res_T = elem(f, i, J);
                        const U nfT   = fxT[eq] * nx + fyT[eq] * ny;
                        for (int l = 0; l < n; ++l) res_T[idx(eq, k, l)] += sc * (Gs[eq] - nfT) * dL[l];
                    }
                }
            }
        }

        // physics_impl_2D.hpp:644-902 residual_kernel (single rank: start=0, max=N)
        //
        // OpenMP: the volume loop writes one element per iteration, an x-edge
        // row j only the elements of row j, a y-edge column i only those of
        // column i; every element still receives its contributions in the
        // serial order (volume, x-edges by I, y-edges by J), so the result is
        // bitwise the single-thread one.
        template <class U>
        void residual(const U* x, U* f, double t) const
        {
            const long Nx = d.nx, Ny = d.ny;
#pragma omp parallel for schedule(static)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx; ++i) volume_term(elem(x, i, j), elem(f, i, j), i, j);
            const bool phys_left = !per_x, phys_right = !per_x;
#pragma omp parallel for schedule(static)
            for (long j = 0; j < Ny; ++j)
            {
                if (phys_left)
                    x_edge(x, f, 0, j, false, true, true, false, t);
                else
                    x_edge(x, f, 0, j, false, true, false, false, t);
                for (long I = 1; I < Nx; ++I) x_edge(x, f, I, j, true, true, false, false, t);
                if (phys_right)
                    x_edge(x, f, Nx, j, true, false, false, true, t);
                else
                    x_edge(x, f, Nx, j, true, false, false, false, t);
            }
            const bool phys_bot = !per_y, phys_top = !per_y;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < Nx; ++i)
            {
                if (phys_bot)
                    y_edge(x, f, i, 0, false, true, true, false, t);
                else
                    y_edge(x, f, i, 0, false, true, false, false, t);
                for (long J = 1; J < Ny; ++J) y_edge(x, f, i, J, true, true, false, false, t);
                if (phys_top)
                    y_edge(x, f, i, Ny, true, false, false, true, t);
                else
                    y_edge(x, f, i, Ny, true, false, false, false, t);
            }
        }

        // physics_impl_2D.hpp:76-124 limit_solution: x-lines (rows) of the
        // element and its left/right neighbours -> scratch, y-lines (columns)
        // with the bottom/top neighbours -> out, out = (scratch + out) / 2.
        // Beyond a non-periodic boundary the neighbour is the ghost layer of
        // the DM_BOUNDARY_GHOSTED local vector (petsc_solver_impl.hpp:502-512,
        // vec_dof_view.hpp:52-60), which DMGlobalToLocal never writes: zeros.
        void filter(const double* sol, double* fil) const
        {
            const long       nd = ndofs_elem;
            const Stabilizer st(tb, d.apply_filter != 0, d.apply_weno != 0, d.weno_gamma, d.shock_theta,
                                d.shock_threshold);
            std::vector<double> zero(nd, 0.0);
            auto nb = [&](long i, long j) -> const double*
            {
                if ((i < 0 || i >= d.nx) && !per_x) return zero.data();
                if ((j < 0 || j >= d.ny) && !per_y) return zero.data();
                return elem(sol, i, j);
            };
#pragma omp parallel for schedule(static)
            for (long j = 0; j < d.ny; ++j)
                for (long i = 0; i < d.nx; ++i)
                {
                    double a[9], b[9], c[9], o[9];
                    const double* uL = nb(i - 1, j);
                    const double* uC = elem(sol, i, j);
                    const double* uR = nb(i + 1, j);
                    const double* uB = nb(i, j - 1);
                    const double* uT = nb(i, j + 1);
                    double*       out = elem(fil, i, j);
                    std::vector<double> scratch(nd);
                    for (int eq = 0; eq < 4; ++eq)
                        for (int l = 0; l < n; ++l)
                        {
                            for (int k = 0; k < n; ++k)
                            {
                                a[k] = uL[idx(eq, k, l)];
                                b[k] = uC[idx(eq, k, l)];
                                c[k] = uR[idx(eq, k, l)];
                            }
                            st.apply(a, b, c, o);
                            for (int k = 0; k < n; ++k) scratch[idx(eq, k, l)] = o[k];
                        }
                    for (int eq = 0; eq < 4; ++eq)
                        for (int k = 0; k < n; ++k)
                        {
                            for (int l = 0; l < n; ++l)
                            {
                                a[l] = uB[idx(eq, k, l)];
                                b[l] = uC[idx(eq, k, l)];
                                c[l] = uT[idx(eq, k, l)];
                            }
                            st.apply(a, b, c, o);
                            for (int l = 0; l < n; ++l) out[idx(eq, k, l)] = o[l];
                        }
                    for (long q = 0; q < nd; ++q) out[q] = 0.5 * (scratch[q] + out[q]);
                }
        }

        // physics_impl_2D.hpp:156-201 impose_ic (Cartesian node positions; curved
        // geometries place the state at the mapped physical coordinates)
        void node_pos(long i, long j, int k, int l, double& xp, double& yp) const
        {
            if (geo.kind == 0)
            {
                const double h_x = d.lx / d.nx, h_y = d.ly / d.ny;
                xp = h_x * (i + 0.5 * (1.0 + tb.csi[k]));
                yp = h_y * (j + 0.5 * (1.0 + tb.csi[l]));
            }
            else
            {
                const auto p = geo.phys(comp_x(i, k), comp_y(j, l));
                xp = p.first;
                yp = p.second;
            }
        }
    };

    // factory.hpp:311-352 exact_vortex_solution (gamma hard-wired 1.4)
    inline St4<double> vortex(double time, double x, double y, const double* prm)
    {
        const double u_0 = prm[0], v_0 = prm[1], beta = prm[2], r_v = prm[3], l = prm[4];
        const double gamma{1.4};
        const double pi{std::numbers::pi};
        double       x_c = l / 2.0 + u_0 * time;
        double       y_c = l / 2.0 + v_0 * time;
        x_c              = x_c - l * std::floor(x_c / l);
        y_c              = y_c - l * std::floor(y_c / l);
        const double dx  = x - x_c;
        const double dy  = y - y_c;
        const double r2  = dx * dx + dy * dy;
        const double f   = 1.0 - (r2 / (r_v * r_v));
        const double ef  = std::exp(f);
        const double rho = std::pow(1.0 - (beta * beta * (gamma - 1.0) * (ef * ef)) / (16.0 * gamma * pi * pi),
                                    1.0 / (gamma - 1.0));
        const double u = u_0 - (beta * ef * dy) / (2.0 * pi * r_v);
        const double v = v_0 + (beta * ef * dx) / (2.0 * pi * r_v);
        const double p = std::pow(rho, gamma);
        const double E = (p / (gamma - 1.0)) + 0.5 * rho * (u * u + v * v);
        return {rho, rho * u, rho * v, E};
    }

    inline St4<double> ic2d(const Gas& g, int kind, const double* prm, double t, double x, double y)
    {
        if (kind == 0)
        {
            const double rho = prm[0], u = prm[1], v = prm[2], p = prm[3];
            return {rho, rho * u, rho * v, g.rhoE(rho, std::sqrt(u * u + v * v), p)};
        }
        if (kind == 1) return vortex(t, x, y, prm);
        if (kind == 4)
        {
            // smooth pressure pulse at rest (test state): rho = rho0 (1 + A g), p = p0 (1 + A g)^gamma,
            // g = exp(-((x-xc)^2 + (y-yc)^2) / w^2); prm = {A, w, xc, yc, rho0, p0}
            const double dx = x - prm[2], dy = y - prm[3];
            const double gg = std::exp(-(dx * dx + dy * dy) / (prm[1] * prm[1]));
            const double rho = prm[4] * (1.0 + prm[0] * gg);
            const double pr  = prm[5] * std::pow(1.0 + prm[0] * gg, g.gamma);
            return {rho, 0.0, 0.0, g.rhoE(rho, 0.0, pr)};
        }
        // factory.hpp:423-454 circular explosion; prm[7] > 0 smooths the jump over
        // a tanh layer of that width (test/benchmark variant, not in the reference)
        const double dx = x - prm[5], dy = y - prm[6];
        const double r  = std::sqrt(dx * dx + dy * dy);
        double       rho, pres;
        if (prm[7] > 0.0)
        {
            const double s = 0.5 * (1.0 - std::tanh((r - prm[4]) / prm[7]));
            rho  = prm[2] + (prm[0] - prm[2]) * s;
            pres = prm[3] + (prm[1] - prm[3]) * s;
        }
        else if (r < prm[4])
        {
            rho  = prm[0];
            pres = prm[1];
        }
        else
        {
            rho  = prm[2];
            pres = prm[3];
        }
        const double u = 0.0, v = 0.0, vel = u + v;
        return {rho, rho * u, rho * v, g.rhoE(rho, vel, pres)};
    }

    // ------------------------------------------------------------------
    // 1D evaluator — physics_impl_1D.hpp
    // ------------------------------------------------------------------
    struct Solver1D
    {
        fro_desc  d;
        GLLTables tb;
        Gas       gas;
        int       n{0};
        bool      periodic{false};
        int       bc[2]{1, 1};
        double    bcs[2][3]{};

        explicit Solver1D(const fro_desc& dd) : d(dd)
        {
            tb       = make_tables(d.p, d.c_vcjh, d.filt_cutoff, d.filt_strength, d.filt_alpha);
            n        = tb.n;
            gas      = Gas{d.gas_kind, d.gamma, d.R_spec, d.vdw_a, d.vdw_b};
            periodic = d.bc_type[0] == BC_PERIODIC && d.bc_type[1] == BC_PERIODIC;
            for (int s = 0; s < 2; ++s)
            {
                bc[s] = d.bc_type[s];
                for (int q = 0; q < 3; ++q) bcs[s][q] = d.bc_state[s][q];
            }
        }
        long ndofs() const { return 3L * n * d.nx; }
        double inv_jac() const { return 2.0 / (d.lx / d.nx); } // physics_evaluator.hpp:133-134

        template <class U>
        std::array<U, 3> node(const U* sp, int i) const
        {
            return {sp[0 * n + i], sp[1 * n + i], sp[2 * n + i]};
        }
        // boundary_conditions.hpp:59-139 (1D native Wall / Outflow; fixed Dirichlet)
        template <class U>
        std::array<U, 3> bc_state(int side, const std::array<U, 3>& in) const
        {
            if (bc[side] == BC_WALL) return {in[0], -in[1], in[2]};
            if (bc[side] == BC_DIRICHLET) return {U{bcs[side][0]}, U{bcs[side][1]}, U{bcs[side][2]}};
            return in;
        }
        template <class U>
        std::array<U, 3> riem(const std::array<U, 3>& a, const std::array<U, 3>& b) const
        {
            return d.riemann == 0 ? rusanov1(a, b, gas) : hllc1(a, b, gas);
        }
        // physics_impl_1D.hpp:123-181 local_steady_res
        template <class U>
        void local(const U* e_u, const std::array<U, 3>& prev, const std::array<U, 3>& next,
                   U* res) const
        {
            const auto&  D  = tb.D;
            const double ij = inv_jac();
            for (int q = 0; q < 3 * n; ++q) res[q] = U{};
            for (int j = 0; j < n; ++j)
            {
                const auto Fj = euler_flux1(node(e_u, j), gas);
                for (int i = 0; i < n; ++i)
                {
                    const U Dij = ij * D[i * n + j];
                    res[0 * n + i] += Dij * Fj[0];
                    res[1 * n + i] += Dij * Fj[1];
                    res[2 * n + i] += Dij * Fj[2];
                }
            }
            const auto FdL = euler_flux1(node(e_u, 0), gas);
            const auto FdR = euler_flux1(node(e_u, n - 1), gas);
            const auto FsL = riem(prev, node(e_u, 0));
            const auto FsR = riem(node(e_u, n - 1), next);
            for (int eq = 0; eq < 3; ++eq)
            {
                const U dlt_L = FsL[eq] - FdL[eq];
                const U dlt_R = FsR[eq] - FdR[eq];
                for (int i = 0; i < n; ++i) res[eq * n + i] += ij * (dlt_L * tb.dL[i] + dlt_R * tb.dR[i]);
            }
        }
        // physics_impl_1D.hpp:8-94 residual_kernel (single rank)
        template <class U>
        void residual(const U* x, U* f, double) const
        {
            const long N = d.nx, nd = 3L * n;
            auto sp = [&](long e) { return x + ((e % N + N) % N) * nd; };
            long first = 0, last = N;
            if (!periodic)
            {
                --last;
                {
                    const long e = N - 1;
                    local(sp(e), node(sp(e - 1), n - 1), bc_state(1, node(sp(e), n - 1)), f + e * nd);
                }
                ++first;
                {
                    const long e = 0;
                    local(sp(e), bc_state(0, node(sp(e), 0)), node(sp(e + 1), 0), f + e * nd);
                }
            }
            for (long e = first; e < last; ++e)
            {
                local(sp(e), node(sp(e - 1), n - 1), node(sp(e + 1), 0), f + e * nd);
                if (d.coord_sys == 1) // radial source :70-90, interior loop only
                {
                    const double h  = d.lx / d.nx;
                    const U*     s_ = sp(e);
                    U*           r_ = f + e * nd;
                    for (int m = 0; m < n; ++m)
                    {
                        const double r     = (static_cast<double>(e) + 0.5 * (tb.csi[m] + 1.0)) * h;
                        const double inv_r = (r > 1e-8) ? 1.0 / r : 0.0;
                        const U      rho = s_[m], rhoU = s_[n + m], rhoE = s_[2 * n + m];
                        const U      vel  = rhoU / rho;
                        const U      pres = gas.pressure(rho, rhoE / rho - U(0.5) * vel * vel);
                        r_[m] += inv_r * rhoU;
                        r_[n + m] += inv_r * rhoU * vel;
                        r_[2 * n + m] += inv_r * vel * (rhoE + pres);
                    }
                }
            }
        }
        // physics_impl_1D.hpp:96-121 limit_solution (neighbours beyond a
        // non-periodic end: the zero ghost layer, as in 2D)
        void filter(const double* sol, double* fil) const
        {
            const long       nd = 3L * n;
            const Stabilizer st(tb, d.apply_filter != 0, d.apply_weno != 0, d.weno_gamma, d.shock_theta,
                                d.shock_threshold);
            const bool          per = periodic;
            std::vector<double> zero(nd, 0.0);
            auto nb = [&](long e) -> const double*
            {
                if (e < 0) return per ? sol + (d.nx - 1) * nd : zero.data();
                if (e >= d.nx) return per ? sol : zero.data();
                return sol + e * nd;
            };
            for (long e = 0; e < d.nx; ++e)
                for (int eq = 0; eq < 3; ++eq)
                    st.apply(nb(e - 1) + eq * n, sol + e * nd + eq * n, nb(e + 1) + eq * n, fil + e * nd + eq * n);
        }
    };

    struct Handle
    {
        fro_desc                  d;
        std::unique_ptr<Solver2D> s2;
        std::unique_ptr<Solver1D> s1;
        long ndofs() const { return s2 ? s2->ndofs() : s1->ndofs(); }
        template <class U>
        void residual(const U* x, U* f, double t) const
        {
            if (s2)
                s2->residual(x, f, t);
            else
                s1->residual(x, f, t);
        }
        void filter(const double* u, double* o) const
        {
            if (s2)
                s2->filter(u, o);
            else
                s1->filter(u, o);
        }
        const GLLTables& tb() const { return s2 ? s2->tb : s1->tb; }
    };
} // namespace fro

using namespace fro;

extern "C" {

void* fro_create(const fro_desc* d)
{
    try
    {
        auto* h = new Handle{};
        h->d    = *d;
        if (d->dims == 2)
            h->s2 = std::make_unique<Solver2D>(*d);
        else
            h->s1 = std::make_unique<Solver1D>(*d);
        return h;
    }
    catch (...)
    {
        return nullptr;
    }
}
void fro_destroy(void* h) { delete static_cast<Handle*>(h); }

// Threads of the OpenMP loops (residual, filter, SSP-RK stage combinations);
// returns the previous setting.  Results do not depend on it.
int fro_set_threads(int nthreads)
{
    const int prev = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return prev;
}
long fro_ndofs(void* h) { return static_cast<Handle*>(h)->ndofs(); }

void fro_tables(void* hh, double* csi, double* w, double* D, double* dL, double* dR, double* filt)
{
    const auto& t = static_cast<Handle*>(hh)->tb();
    const int   n = t.n;
    for (int i = 0; i < n; ++i)
    {
        csi[i] = t.csi[i];
        w[i]   = t.w[i];
        dL[i]  = t.dL[i];
        dR[i]  = t.dR[i];
    }
    for (int i = 0; i < n * n; ++i)
    {
        D[i]    = t.D[i];
        filt[i] = t.filt[i];
    }
}

int fro_residual(void* hh, double t, const double* u, double* R)
{
    static_cast<Handle*>(hh)->residual(u, R, t);
    return 0;
}

int fro_residual_cplx(void* hh, double t, const double* ure, const double* uim, double* Rre,
                      double* Rim)
{
    auto*             h = static_cast<Handle*>(hh);
    const long        N = h->ndofs();
    std::vector<cplx> x(N), f(N);
    for (long q = 0; q < N; ++q) x[q] = cplx(ure[q], uim[q]);
    h->residual(x.data(), f.data(), t);
    for (long q = 0; q < N; ++q)
    {
        Rre[q] = f[q].real();
        Rim[q] = f[q].imag();
    }
    return 0;
}

// petsc_solver_impl.hpp:270-359 matmult_2d_impl / :361-432 matmult_1d_impl
} // extern "C"

// prim_to_cons_elem (problem_def.hpp:36-50) over every node; W = (log rho, u, [v,]
// log e) per node in the state layout
template <class U>
static void prim_to_cons_all(const Handle& h, const U* w, U* c)
{
    const int n = h.tb().n;
    if (h.s2)
    {
        const long nodes = h.ndofs() / 4;
        for (long i = 0; i < nodes; ++i)
        {
            const U* a = w + 4 * i;
            U*       b = c + 4 * i;
            const U  rho = std::exp(a[0]);
            const U  kin = U(0.5) * ((U(0.0) + a[1] * a[1]) + a[2] * a[2]);
            const U  E   = rho * (std::exp(a[3]) + kin);
            b[0] = rho;
            b[1] = rho * a[1];
            b[2] = rho * a[2];
            b[3] = E;
        }
        return;
    }
    const long ne = h.ndofs() / (3L * n);
    for (long e = 0; e < ne; ++e)
        for (int m = 0; m < n; ++m)
        {
            const long i0 = (e * 3 + 0) * n + m, i1 = (e * 3 + 1) * n + m, i2 = (e * 3 + 2) * n + m;
            const U    rho = std::exp(w[i0]);
            const U    kin = U(0.5) * (U(0.0) + w[i1] * w[i1]);
            const U    E   = rho * (std::exp(w[i2]) + kin);
            c[i0] = rho;
            c[i1] = rho * w[i1];
            c[i2] = E;
        }
}

// cons_to_prim_elem (problem_def.hpp:55-70)
static void cons_to_prim_all(const Handle& h, const double* c, double* w)
{
    const int n = h.tb().n;
    if (h.s2)
    {
        const long nodes = h.ndofs() / 4;
        for (long i = 0; i < nodes; ++i)
        {
            const double* a = c + 4 * i;
            const double  inv = 1.0 / a[0];
            const double  u = a[1] * inv, v = a[2] * inv;
            const double  kin = 0.5 * ((0.0 + u * u) + v * v);
            const double  r0 = std::log(a[0]), r3 = std::log(a[3] * inv - kin);
            w[4 * i] = r0; w[4 * i + 1] = u; w[4 * i + 2] = v; w[4 * i + 3] = r3;
        }
        return;
    }
    const long ne = h.ndofs() / (3L * n);
    for (long e = 0; e < ne; ++e)
        for (int m = 0; m < n; ++m)
        {
            const long   i0 = (e * 3 + 0) * n + m, i1 = (e * 3 + 1) * n + m, i2 = (e * 3 + 2) * n + m;
            const double inv = 1.0 / c[i0];
            const double u   = c[i1] * inv;
            const double kin = 0.5 * (0.0 + u * u);
            const double r0 = std::log(c[i0]), r2 = std::log(c[i2] * inv - kin);
            w[i0] = r0; w[i1] = u; w[i2] = r2;
        }
}

extern "C" {

int fro_prim_to_cons(void* hh, const double* w, double* c)
{
    prim_to_cons_all(*static_cast<Handle*>(hh), w, c);
    return 0;
}

int fro_cons_to_prim(void* hh, const double* c, double* w)
{
    cons_to_prim_all(*static_cast<Handle*>(hh), c, w);
    return 0;
}

// fun_implicit (petsc_solver_impl.hpp:626-655)
int fro_ifunction(void* hh, double t, const double* u, const double* udot, double* F)
{
    auto*               h = static_cast<Handle*>(hh);
    const long          N = h->ndofs();
    std::vector<double> cu(u, u + N);
    if (h->d.cond_sys) prim_to_cons_all(*h, u, cu.data());
    h->residual(cu.data(), F, t);
    for (long q = 0; q < N; ++q) F[q] = udot[q] + F[q];
    return 0;
}

int fro_jvp_cstep(void* hh, double t, const double* u, const double* v, double sigma, double hs,
                  double* out)
{
    auto*             h          = static_cast<Handle*>(hh);
    const long        N          = h->ndofs();
    const double      one_over_h = 1.0 / hs;
    std::vector<cplx> uiv(N), f(N);
    for (long q = 0; q < N; ++q) uiv[q] = cplx(u[q], v[q] * hs);
    if (h->d.cond_sys) // Q = M(u + i v h): prim-to-cons in complex arithmetic (:297-326, :386-401)
    {
        std::vector<cplx> w(uiv);
        prim_to_cons_all(*h, w.data(), uiv.data());
    }
    for (long q = 0; q < N; ++q) out[q] = sigma * std::imag(uiv[q]) * one_over_h;
    h->residual(uiv.data(), f.data(), t);
    for (long q = 0; q < N; ++q) out[q] += std::imag(f[q]) * one_over_h;
    return 0;
}

int fro_filter(void* hh, const double* u, double* out)
{
    static_cast<Handle*>(hh)->filter(u, out);
    return 0;
}

// PETSc TSSSP (rks3) stage combinations with rhs = -R (petsc_solver_impl.hpp:657-670),
// post-step filter (petsc_solver_impl.hpp:604-605, 683-702).
int fro_ssprk_step(void* hh, double t, double dt, double* u, int stages)
{
    auto*               h = static_cast<Handle*>(hh);
    const long          N = h->ndofs();
    std::vector<double> u0(u, u + N), a(N), b(N), R(N);
    auto L = [&](const double* x, double tt, double* out)
    {
        h->residual(x, out, tt);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) out[q] = -out[q];
    };
    if (stages == 3)
    {
        L(u0.data(), t, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) a[q] = u0[q] + dt * R[q];
        L(a.data(), t + dt, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) b[q] = 0.75 * u0[q] + 0.25 * (a[q] + dt * R[q]);
        L(b.data(), t + 0.5 * dt, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) u[q] = (1.0 / 3.0) * u0[q] + (2.0 / 3.0) * (b[q] + dt * R[q]);
    }
    else if (stages == 4)
    {
        const double hdt = 0.5 * dt;
        L(u0.data(), t, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) a[q] = u0[q] + hdt * R[q];
        L(a.data(), t + hdt, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) b[q] = a[q] + hdt * R[q];
        L(b.data(), t + dt, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) a[q] = (2.0 / 3.0) * u0[q] + (1.0 / 3.0) * (b[q] + hdt * R[q]);
        L(a.data(), t + hdt, R.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < N; ++q) u[q] = a[q] + hdt * R[q];
    }
    else
        return 1;
    if (h->d.apply_filter || h->d.apply_weno) // filters_active() (filter.hpp:710)
    {
        std::vector<double> tmp(u, u + N);
        h->filter(tmp.data(), u);
    }
    return 0;
}

static int impose_ic_cons(void* hh, int kind, const double* prm, double* u)
{
    auto* h = static_cast<Handle*>(hh);
    if (h->s2)
    {
        const auto& s = *h->s2;
#pragma omp parallel for schedule(static)
        for (long j = 0; j < s.d.ny; ++j)
            for (long i = 0; i < s.d.nx; ++i)
            {
                double* e = u + (j * s.d.nx + i) * s.ndofs_elem;
                for (int k = 0; k < s.n; ++k)
                    for (int l = 0; l < s.n; ++l)
                    {
                        double xp, yp;
                        s.node_pos(i, j, k, l, xp, yp);
                        const auto st = ic2d(s.gas, kind, prm, 0.0, xp, yp);
                        for (int eq = 0; eq < 4; ++eq) e[s.idx(eq, k, l)] = st[eq];
                    }
            }
        return 0;
    }
    // physics_impl_1D.hpp:207-244 with factory.hpp:196-214 (Sod) / uniform
    const auto&  s   = *h->s1;
    const double hh_ = s.d.lx / s.d.nx;
    for (long e = 0; e < s.d.nx; ++e)
        for (int q = 0; q < s.n; ++q)
        {
            const double pos = hh_ * (e + 0.5 * (1.0 + s.tb.csi[q]));
            double       rho, vel, pres;
            if (kind == 3)
            {
                const double x_c = s.d.lx / 2.0;
                if (pos < x_c)
                {
                    rho = prm[0]; vel = prm[1]; pres = prm[2];
                }
                else
                {
                    rho = prm[3]; vel = prm[4]; pres = prm[5];
                }
            }
            else
            {
                rho = prm[0]; vel = prm[1]; pres = prm[3];
            }
            double* sp        = u + e * 3L * s.n;
            sp[0 * s.n + q]   = rho;
            sp[1 * s.n + q]   = rho * vel;
            sp[2 * s.n + q]   = s.gas.rhoE(rho, vel, pres);
        }
    return 0;
}

// impose_ic (physics_impl_2D.hpp:156-201, physics_impl_1D.hpp:207-244): with
// cond_sys the state is stored as (log rho, u, [v,] log e)
int fro_impose_ic(void* hh, int kind, const double* prm, double* u)
{
    const int rc = impose_ic_cons(hh, kind, prm, u);
    auto*     h  = static_cast<Handle*>(hh);
    if (rc == 0 && h->d.cond_sys) cons_to_prim_all(*h, u, u);
    return rc;
}

// physics_impl_2D.hpp:203-265 compute_error (exact_sol = same kinds as the ICs)
int fro_l2_error(void* hh, int kind, const double* prm, double t, const double* u, double* out)
{
    auto* h = static_cast<Handle*>(hh);
    if (!h->s2) return 1;
    const auto&         s = *h->s2;
    std::vector<double> cu;
    if (h->d.cond_sys) // compute_error converts the stored primitives (physics_impl_2D.hpp:229-240)
    {
        cu.resize(h->ndofs());
        prim_to_cons_all(*h, u, cu.data());
        u = cu.data();
    }
    double errs[4]{};
    for (long j = 0; j < s.d.ny; ++j)
        for (long i = 0; i < s.d.nx; ++i)
        {
            const double* e = u + (j * s.d.nx + i) * s.ndofs_elem;
            for (int k = 0; k < s.n; ++k)
                for (int l = 0; l < s.n; ++l)
                {
                    const auto [xp, yp] = s.geo.phys(s.comp_x(i, k), s.comp_y(j, l));
                    const auto m        = s.metrics(s.comp_x(i, k), s.comp_y(j, l));
                    const auto st       = ic2d(s.gas, kind, prm, t, xp, yp);
                    const double cv     = s.tb.w[k] * s.tb.w[l] * m.detJ;
                    for (int eq = 0; eq < 4; ++eq)
                    {
                        const double dd = e[s.idx(eq, k, l)] - st[eq];
                        errs[eq] += cv * dd * dd;
                    }
                }
        }
    for (int eq = 0; eq < 4; ++eq) out[eq] = std::sqrt(errs[eq]);
    return 0;
}

int fro_flux2d(void* hh, const double* U, double* F, double* G)
{
    auto*       h = static_cast<Handle*>(hh);
    St4<double> s{U[0], U[1], U[2], U[3]}, f, g;
    double      p;
    euler_fluxes_2d_w_p(s, h->s2->gas, f, g, p);
    for (int q = 0; q < 4; ++q)
    {
        F[q] = f[q];
        G[q] = g[q];
    }
    return 0;
}
int fro_riemann2d(void* hh, const double* UL, const double* UR, double nx, double ny, double* out)
{
    auto*       h = static_cast<Handle*>(hh);
    St4<double> a{UL[0], UL[1], UL[2], UL[3]}, b{UR[0], UR[1], UR[2], UR[3]};
    const auto  r = h->s2->riemann(a, b, nx, ny);
    for (int q = 0; q < 4; ++q) out[q] = r[q];
    return 0;
}
int fro_riemann1d(void* hh, const double* UL, const double* UR, double* out)
{
    auto*                 h = static_cast<Handle*>(hh);
    std::array<double, 3> a{UL[0], UL[1], UL[2]}, b{UR[0], UR[1], UR[2]};
    const auto            r = h->s1->riem(a, b);
    for (int q = 0; q < 3; ++q) out[q] = r[q];
    return 0;
}
int fro_metrics(void* hh, double cx, double cy, double* out)
{
    auto*      h  = static_cast<Handle*>(hh);
    const auto m  = h->s2->metrics(cx, cy);
    const auto fx = h->s2->geo.x_face(cx, cy);
    const auto fy = h->s2->geo.y_face(cx, cy);
    out[0] = m.xi_x; out[1] = m.xi_y; out[2] = m.eta_x; out[3] = m.eta_y; out[4] = m.detJ;
    out[5] = fx.nx; out[6] = fx.ny; out[7] = fy.nx; out[8] = fy.ny;
    return 0;
}

} // extern "C"

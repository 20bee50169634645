// san_main.cpp — driver of the sanitised oracle build (test infrastructure
// only): runs every C entry point of fr_oracle.h on small 2D and 1D problems
// so that AddressSanitizer and UndefinedBehaviorSanitizer see all code paths
// (curved maps, every boundary kind, ideal and van der Waals gas, Rusanov and
// HLLC, filter and WENO limiter, complex step, log variables, radial source).
// Exit status 0 and no sanitizer report = pass (UBSan does not recover).
#include "fr_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace
{
    fro_desc base2d(int nx, int ny, int p)
    {
        fro_desc d{};
        d.dims = 2;
        d.nx = nx;
        d.ny = ny;
        d.p = p;
        d.lx = d.ly = 2.0;
        d.gamma = 1.4;
        d.R_spec = 287.05;
        d.riemann = 1;
        for (int k = 0; k < 4; ++k) d.bc_type[k] = 1;
        d.filt_cutoff = 0;
        d.filt_strength = 16;
        d.filt_alpha = 36.0;
        d.weno_gamma = 0.998;
        d.shock_theta = 1.0;
        d.shock_threshold = 0.8;
        return d;
    }

    int run(const char* name, const fro_desc& d, int ic_kind, const std::vector<double>& prm)
    {
        void* h = fro_create(&d);
        if (!h) { std::printf("%s: create failed\n", name); return 1; }
        const long N = fro_ndofs(h);
        std::vector<double> u(N), R(N), im(N, 0.0), Ri(N), v(N), out(N), w(N), c(N);
        const int n = d.p + 1;
        std::vector<double> csi(n), wq(n), D(n * n), dL(n), dR(n), F(n * n);
        fro_tables(h, csi.data(), wq.data(), D.data(), dL.data(), dR.data(), F.data());
        std::vector<double> pr(prm);
        pr.resize(8, 0.0);
        fro_impose_ic(h, ic_kind, pr.data(), u.data());
        for (long i = 0; i < N; ++i)
        {
            v[i]  = std::sin(0.37 * i);
            im[i] = 1e-7 * v[i];
        }
        fro_residual(h, 0.0, u.data(), R.data());
        fro_residual_cplx(h, 0.0, u.data(), im.data(), out.data(), Ri.data());
        fro_jvp_cstep(h, 0.0, u.data(), v.data(), 1.0, 1e-7, out.data());
        fro_ifunction(h, 0.0, u.data(), v.data(), out.data());
        fro_filter(h, u.data(), out.data());
        double err[4];
        fro_l2_error(h, ic_kind, pr.data(), 0.0, u.data(), err);
        fro_cons_to_prim(h, u.data(), w.data());
        fro_prim_to_cons(h, w.data(), c.data());
        std::vector<double> s(u);
        fro_ssprk_step(h, 0.0, 1e-5, s.data(), 3);
        fro_ssprk_step(h, 0.0, 1e-5, s.data(), 4);
        double sum = 0.0;
        for (long i = 0; i < N; ++i) sum += std::fabs(R[i]) + std::fabs(s[i]);
        fro_destroy(h);
        // (the log-variable case feeds primitives to the explicit residual: its
        // sum need not be finite; the sanitizers are the check here)
        std::printf("%-32s ndofs %8ld  sum %.6e\n", name, N, sum);
        return 0;
    }
} // namespace

int main()
{
    int bad = 0;
    const std::vector<double> pulse{0.2, 0.5, 1.0, 1.0, 1.0, 1.0};
    {
        fro_desc d = base2d(5, 4, 3);
        d.bc_type[0] = d.bc_type[1] = d.bc_type[2] = d.bc_type[3] = 5;
        bad |= run("periodic_hllc_ideal_p3", d, 4, pulse);
    }
    {
        fro_desc d = base2d(4, 3, 7);
        d.geometry = 4;
        d.geom_param[0] = 0.05;
        d.geom_param[1] = d.geom_param[2] = -1.0;
        d.gas_kind = 1;
        d.vdw_a = 0.1;
        d.vdw_b = 0.02;
        d.apply_filter = 1;
        bad |= run("warped_vdw_filter_p7", d, 4, {0.2, 0.5, 0.0, 0.0, 1.0, 1.0});
    }
    {
        fro_desc d = base2d(6, 5, 2);
        d.riemann = 0;
        d.apply_weno = 1;
        d.bc_type[0] = d.bc_type[1] = d.bc_type[2] = d.bc_type[3] = 0; // walls
        bad |= run("walls_rusanov_weno_p2", d, 2, {1.0, 1.0, 0.125, 0.1, 0.4, 1.0, 1.0});
    }
    {
        fro_desc d = base2d(6, 3, 4);
        d.lx = 3.0;
        d.ly = 0.8;
        d.geometry = 3;
        d.geom_param[0] = 0.0625;
        d.geom_param[1] = -1.5;
        d.geom_param[3] = 1.0;
        const double g = 1.4, uin = 0.5 * std::sqrt(g);
        const double inflow[4] = {1.0, uin, 0.0, 1.0 / (g - 1) + 0.5 * uin * uin};
        d.bc_type[0] = 2;
        d.bc_type[1] = 1;
        d.bc_type[2] = 2;
        d.bc_type[3] = 6;
        std::memcpy(d.bc_state[0], inflow, sizeof(inflow));
        std::memcpy(d.bc_state[2], inflow, sizeof(inflow));
        d.cond_sys = 1;
        d.coord_sys = 1;
        bad |= run("bump_dirichlet_slipwall_logvar", d, 0, {1.0, uin, 0.0, 1.0});
    }
    {
        fro_desc d{};
        d.dims = 1;
        d.nx = 30;
        d.ny = 1;
        d.p = 3;
        d.lx = 1.0;
        d.gamma = 1.4;
        d.riemann = 1;
        d.bc_type[0] = d.bc_type[1] = 1;
        d.apply_filter = 1;
        d.filt_cutoff = 0;
        d.filt_strength = 8;
        d.filt_alpha = 36.0;
        d.weno_gamma = 0.998;
        d.shock_theta = 1.0;
        d.shock_threshold = 0.8;
        bad |= run("sod1d_hllc_filter_p3", d, 3, {1.0, 0.0, 1.0, 0.125, 0.0, 0.1});
    }
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad;
}

// frsdg_geometry.hpp — analytic maps (csi,eta) in [0,1]^2 -> (x,y) and their
// Jacobians, evaluated once per context on the GPU to fill the metric tables
// in HBM (the reference recomputes them per node per call,
// physics_impl_2D.hpp:327, 704-705, 812-813).
//
// Models (same maps as the reference's geometry policies):
//   CARTESIAN  geometry_models.hpp:72-107         x = Lx csi, y = Ly eta
//   DELAVAL    geometry_models.hpp:127-195        straight profile h = 0.5;
//              the metric assumes y = h eta while the coordinates use
//              y = h(2 eta - 1) — kept as in the reference
//   WEDGE      tests/bump.cpp:13-171              curved wedge wall
//   BUMP       channel bump y_w = A sin^2(pi (x - xc + w/2)/w), |x - xc| <= w/2,
//              vertical blending y = y_w (1 - eta) + Ly eta (bump.cpp pattern)
//   WARPED     x = X + A sin(pi X) sin(pi Y), y = Y + A sin(pi X) sin(pi Y),
//              X = x0 + Lx csi, Y = y0 + Ly eta
// Face normals are the unit contravariant directions grad(xi)/|grad(xi)| and
// grad(eta)/|grad(eta)|, which reproduce the models' own normals (Cartesian
// (1,0)/(0,1); wedge/bump x-faces (1,0) and y-faces (-y_csi, x_csi)/norm).
#pragma once
#include "frsdg_scalar.hpp"

namespace frsdg
{
    struct GeomP
    {
        int    kind;
        double Lx, Ly;
        double prm[4];
    };
    struct Jac2
    {
        double x_c, y_c, x_e, y_e; // dx/dcsi, dy/dcsi, dx/deta, dy/deta
    };

    FR_HD void wedge_profile(double Lx, double x, double& h, double& dh)
    {
        const double x_start = 0.15 * Lx, x_c = 0.30 * Lx, x_es = 0.55 * Lx, x_ee = 0.70 * Lx;
        const double cs = 0.2, es = -0.2;
        const double A_c = cs / (2.0 * (x_c - x_start));
        const double dxc = x_c - x_start;
        const double hce = A_c * dxc * dxc;
        if (x <= x_start)
        {
            h = 0.0;
            dh = 0.0;
            return;
        }
        if (x <= x_c)
        {
            const double d = x - x_start;
            h              = -(A_c * d * d);
            dh             = -(2.0 * A_c * d);
            return;
        }
        if (x <= x_es)
        {
            const double d = x - x_c;
            h              = -(hce + cs * d);
            dh             = -cs;
            return;
        }
        const double hes = hce + cs * (x_es - x_c);
        const double dxe = x_ee - x_es;
        const double A_e = (es - cs) / (2.0 * dxe);
        if (x <= x_ee)
        {
            const double d = x - x_es;
            h              = -(hes + cs * d + A_e * d * d);
            dh             = -(cs + 2.0 * A_e * d);
            return;
        }
        const double hee = hes + cs * dxe + A_e * dxe * dxe;
        h                = -(hee + es * (x - x_ee));
        dh               = -es;
    }
    FR_HD void bump_profile(const double* prm, double x, double& h, double& dh)
    {
        const double A = prm[0], xc = prm[2], w = prm[3];
        if (::fabs(x - xc) > 0.5 * w)
        {
            h  = 0.0;
            dh = 0.0;
            return;
        }
        const double pi = 3.141592653589793238462643383279502884;
        const double s  = pi * (x - xc + 0.5 * w) / w;
        const double ss = ::sin(s), cc = ::cos(s);
        h               = A * ss * ss;
        dh              = A * 2.0 * ss * cc * pi / w;
    }

    FR_HD void geom_phys(const GeomP& g, double csi, double eta, double& x, double& y)
    {
        const double pi = 3.141592653589793238462643383279502884;
        switch (g.kind)
        {
            case 0:
                x = g.Lx * csi;
                y = g.Ly * eta;
                return;
            case 1:
                x = g.Lx * csi;
                y = -0.5 * (1.0 - eta) + 0.5 * eta;
                return;
            case 2:
            {
                double h, dh;
                x = g.Lx * csi;
                wedge_profile(g.Lx, x, h, dh);
                y = h * (1.0 - eta) + g.Ly * eta;
                return;
            }
            case 3:
            {
                double h, dh;
                x = g.prm[1] + g.Lx * csi;
                bump_profile(g.prm, x, h, dh);
                y = h * (1.0 - eta) + g.Ly * eta;
                return;
            }
            default:
            {
                const double X = g.prm[1] + g.Lx * csi, Y = g.prm[2] + g.Ly * eta;
                const double s = ::sin(pi * X) * ::sin(pi * Y);
                x              = X + g.prm[0] * s;
                y              = Y + g.prm[0] * s;
                return;
            }
        }
    }
    FR_HD Jac2 geom_jac(const GeomP& g, double csi, double eta)
    {
        const double pi = 3.141592653589793238462643383279502884;
        switch (g.kind)
        {
            case 0: return {g.Lx, 0.0, 0.0, g.Ly};
            case 1: return {g.Lx, 0.0, 0.0, 0.5};
            case 2:
            {
                double h, dh;
                wedge_profile(g.Lx, g.Lx * csi, h, dh);
                return {g.Lx, dh * g.Lx * (1.0 - eta), 0.0, g.Ly - h};
            }
            case 3:
            {
                double h, dh;
                bump_profile(g.prm, g.prm[1] + g.Lx * csi, h, dh);
                return {g.Lx, dh * g.Lx * (1.0 - eta), 0.0, g.Ly - h};
            }
            default:
            {
                const double A = g.prm[0];
                const double X = g.prm[1] + g.Lx * csi, Y = g.prm[2] + g.Ly * eta;
                const double dsX = pi * ::cos(pi * X) * ::sin(pi * Y);
                const double dsY = pi * ::sin(pi * X) * ::cos(pi * Y);
                return {g.Lx * (1.0 + A * dsX), g.Lx * A * dsX, g.Ly * A * dsY, g.Ly * (1.0 + A * dsY)};
            }
        }
    }

    // Composite contravariant metrics of (r,s) -> (x,y) at computational
    // coordinates (physics_evaluator.hpp:282-305).
    struct VolM
    {
        double xi_x, xi_y, eta_x, eta_y, detJ;
    };
    FR_HD VolM composite_metrics(const GeomP& g, int nx, int ny, double csi, double eta)
    {
        const Jac2   J     = geom_jac(g, csi, eta);
        const double dc_dr = 0.5 / static_cast<double>(nx);
        const double de_ds = 0.5 / static_cast<double>(ny);
        const double x_r = J.x_c * dc_dr, y_r = J.y_c * dc_dr;
        const double x_s = J.x_e * de_ds, y_s = J.y_e * de_ds;
        const double det = x_r * y_s - x_s * y_r;
        return {y_s / det, -x_s / det, -y_r / det, x_r / det, det};
    }
} // namespace frsdg

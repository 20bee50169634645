// frsdg_tables.hpp — host-side setup tables of one polynomial degree, computed
// once per context and uploaded to HBM:
//   GLL nodes   : Newton iteration on (1-x^2) P_p'(x) from Chebyshev-Gauss-
//                 Lobatto starting points (the reference uses the Golub-Welsch
//                 eigenvalues of the Jacobi matrix, reference_element.hpp:91-147;
//                 both give the roots of P_p');
//   GLL weights : w = 2 / (p (p+1) P_p(x)^2)   (reference_element.hpp:104-115);
//   D           : barycentric Lagrange derivative, diagonal by the negative
//                 row sum (the reference builds the same matrix from node
//                 products, cached_computations.hpp:107-138);
//   dL, dR      : VCJH correction-function derivatives at the nodes, the
//                 Vincent-Castonguay-Jameson formula with eta_p =
//                 c (2p+1) (a_p p!)^2 / 2, a_p = (2p)!/(2^p (p!)^2)
//                 (cached_computations.hpp:41-84);
//   filter      : F = V diag(sigma) V^-1, sigma_i = 1 for i < n_c and
//                 exp(-alpha ((i-n_c)/(p-n_c))^s) otherwise (filter.hpp:451-484);
//                 V^-1 by the exact discrete Legendre transform on the GLL
//                 quadrature (V^-1_jk = w_k P_j(x_k) / sum_m w_m P_j(x_m)^2),
//                 instead of a general matrix inverse.
#pragma once
#include <cmath>
#include <vector>

namespace frsdg
{
    struct Tables
    {
        int                 p = 0, n = 0;
        std::vector<double> csi, w, D, dL, dR, filt;
    };

    // P_j(x) for j = 0..m and their derivatives by the Bonnet recurrence
    inline void legendre_all(int m, double x, double* P, double* dP)
    {
        P[0]  = 1.0;
        dP[0] = 0.0;
        if (m >= 1)
        {
            P[1]  = x;
            dP[1] = 1.0;
        }
        for (int j = 1; j < m; ++j)
        {
            P[j + 1]  = ((2.0 * j + 1.0) * x * P[j] - j * P[j - 1]) / (j + 1.0);
            dP[j + 1] = ((2.0 * j + 1.0) * (P[j] + x * dP[j]) - j * dP[j - 1]) / (j + 1.0);
        }
    }

    inline bool make_tables(int p, double c_vcjh, int n_cut, int strength, double alpha, Tables& t)
    {
        if (p < 1 || p > 8) return false;
        const int n = p + 1;
        t.p         = p;
        t.n         = n;
        t.csi.assign(n, 0.0);
        t.w.assign(n, 0.0);
        const double pi = 3.141592653589793238462643383279502884;
        double       P[12], dP[12];
        // --- nodes: roots of (1-x^2) P_p'(x)
        for (int k = 0; k < n; ++k)
        {
            double x = -std::cos(pi * k / p);
            if (k > 0 && k < p)
            {
                for (int it = 0; it < 100; ++it)
                {
                    // f = P_p'(x); f' = P_p''(x) = (2x P_p' - p(p+1) P_p) / (1 - x^2)
                    legendre_all(p, x, P, dP);
                    const double f  = dP[p];
                    const double fp = (2.0 * x * dP[p] - p * (p + 1.0) * P[p]) / (1.0 - x * x);
                    const double dx = f / fp;
                    x -= dx;
                    if (std::fabs(dx) < 1e-17) break;
                }
            }
            t.csi[k] = x;
        }
        // exact symmetry and centre node
        for (int k = 0; k < n / 2; ++k)
        {
            const double s       = 0.5 * (t.csi[n - 1 - k] - t.csi[k]);
            t.csi[k]             = -s;
            t.csi[n - 1 - k]     = s;
        }
        if (n % 2 == 1) t.csi[n / 2] = 0.0;
        for (int k = 0; k < n; ++k)
        {
            legendre_all(p, t.csi[k], P, dP);
            t.w[k] = 2.0 / (p * (p + 1.0) * P[p] * P[p]);
        }
        // --- D by barycentric weights
        std::vector<double> bw(n, 1.0);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k)
                if (k != j) bw[j] /= (t.csi[j] - t.csi[k]);
        t.D.assign(n * n, 0.0);
        for (int i = 0; i < n; ++i)
        {
            double s = 0.0;
            for (int j = 0; j < n; ++j)
            {
                if (j == i) continue;
                const double d = (bw[j] / bw[i]) / (t.csi[i] - t.csi[j]);
                t.D[i * n + j] = d;
                s += d;
            }
            t.D[i * n + i] = -s;
        }
        // --- VCJH correction derivatives
        double fact_p = 1.0, fact_2p = 1.0;
        for (int i = 2; i <= p; ++i) fact_p *= i;
        for (int i = 2; i <= 2 * p; ++i) fact_2p *= i;
        const double a_p   = fact_2p / (std::ldexp(1.0, p) * fact_p * fact_p);
        const double eta_p = c_vcjh * (2.0 * p + 1.0) * (a_p * fact_p) * (a_p * fact_p) / 2.0;
        const double sgn   = (p % 2 == 0) ? 1.0 : -1.0;
        t.dL.assign(n, 0.0);
        t.dR.assign(n, 0.0);
        for (int i = 0; i < n; ++i)
        {
            legendre_all(p + 1, t.csi[i], P, dP);
            const double corr = (eta_p * dP[p - 1] + dP[p + 1]) / (1.0 + eta_p);
            t.dR[i]           = 0.5 * (dP[p] + corr);
            t.dL[i]           = sgn * 0.5 * (dP[p] - corr);
        }
        // --- modal exponential filter
        std::vector<double> V(n * n), Vinv(n * n), sig(n), gam(n, 0.0);
        for (int i = 0; i < n; ++i)
        {
            legendre_all(p, t.csi[i], P, dP);
            for (int j = 0; j < n; ++j) V[i * n + j] = P[j];
        }
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) gam[j] += t.w[k] * V[k * n + j] * V[k * n + j];
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) Vinv[j * n + k] = t.w[k] * V[k * n + j] / gam[j];
        for (int i = 0; i < n; ++i)
        {
            if (i < n_cut)
                sig[i] = 1.0;
            else
            {
                const double r = double(i - n_cut) / double((n - 1) - n_cut);
                sig[i]         = std::exp(-alpha * std::pow(r, double(strength)));
            }
        }
        t.filt.assign(n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
            {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += V[i * n + k] * sig[k] * Vinv[k * n + j];
                t.filt[i * n + j] = s;
            }
        return true;
    }
} // namespace frsdg

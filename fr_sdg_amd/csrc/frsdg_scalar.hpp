// frsdg_scalar.hpp — the three scalar types the residual kernels are
// instantiated on:
//   double  — the plain residual;
//   Dual    — forward tangent (value + one directional derivative); branches
//             are taken on the value (primal) part;
//   Cplx    — the complex step of petsc_solver_impl.hpp:270-359, with the
//             reference's branch rule re+im (riemann_solver.hpp:299-305) and
//             the CSD-safe wrap_positive of macro_and_libc.hpp:115-138.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

#define FR_HD __host__ __device__ __forceinline__

namespace frsdg
{
    struct Dual
    {
        double v, d;
    };
    struct Cplx
    {
        double re, im;
    };

    // Reciprocal and quotient of doubles from the hardware reciprocal estimate,
    // refined by two Newton steps (r <- r + r (1 - x r)) and, for the
    // quotient, one residual correction q <- q + r (a - b q): within one unit
    // in the last place of the correctly rounded result for normal operands,
    // without the range scaling of the IEEE divide sequence (a zero or
    // infinite divisor gives NaN).  Dual forms its value part with them (the
    // same primal as the double kernels); Cplx divides as written.
    FR_HD double rcp_(double x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        double r = __builtin_amdgcn_rcp(x);
        r        = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
        r        = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
        return r;
#else
        return 1.0 / x;
#endif
    }
    FR_HD double quot_(double a, double b)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const double r = rcp_(b);
        const double q = a * r;
        return __builtin_fma(__builtin_fma(-b, q, a), r, q);
#else
        return a / b;
#endif
    }
    // Square root of a double from the hardware reciprocal-square-root
    // estimate y: g = x y, h = y / 2, one Goldschmidt step, then two residual
    // corrections g <- g + h (x - g^2); within one unit in the last place for
    // positive normal x, +-0 kept, NaN for x < 0 (no range scaling: infinite
    // and subnormal x are outside its domain).
    FR_HD double sqrt_(double x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const double y = __builtin_amdgcn_rsq(x);
        double       g = x * y, h = 0.5 * y;
        const double r = __builtin_fma(-g, h, 0.5);
        g              = __builtin_fma(g, r, g);
        h              = __builtin_fma(h, r, h);
        g              = __builtin_fma(__builtin_fma(-g, g, x), h, g);
        g              = __builtin_fma(__builtin_fma(-g, g, x), h, g);
        return (x == 0.0) ? x : g;
#else
        return ::sqrt(x);
#endif
    }
    // ---------------- Dual ----------------
    FR_HD Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
    FR_HD Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
    FR_HD Dual operator-(Dual a) { return {-a.v, -a.d}; }
    FR_HD Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
    FR_HD Dual operator/(Dual a, Dual b)
    {
        const double q = a.v / b.v;
        return {q, (a.d - q * b.d) / b.v};
    }
    FR_HD Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
    FR_HD Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
    FR_HD Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
    FR_HD Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
    FR_HD Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
    FR_HD Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
    FR_HD Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
    FR_HD Dual operator/(double a, Dual b)
    {
        const double q = a / b.v;
        return {q, -q * b.d / b.v};
    }
    FR_HD Dual sqrt(Dual a)
    {
        const double s = sqrt_(a.v);
        return {s, a.d / (2.0 * s)};
    }

    // ---------------- Cplx ----------------
    FR_HD Cplx operator+(Cplx a, Cplx b) { return {a.re + b.re, a.im + b.im}; }
    FR_HD Cplx operator-(Cplx a, Cplx b) { return {a.re - b.re, a.im - b.im}; }
    FR_HD Cplx operator-(Cplx a) { return {-a.re, -a.im}; }
    FR_HD Cplx operator*(Cplx a, Cplx b)
    {
        return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    }
    FR_HD Cplx operator/(Cplx a, Cplx b)
    {
        // Smith's algorithm: scale by the larger component of the divisor.
        if (::fabs(b.re) >= ::fabs(b.im))
        {
            const double r = b.im / b.re, den = b.re + b.im * r;
            return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
        }
        const double r = b.re / b.im, den = b.re * r + b.im;
        return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
    }
    FR_HD Cplx operator+(Cplx a, double b) { return {a.re + b, a.im}; }
    FR_HD Cplx operator+(double a, Cplx b) { return {a + b.re, b.im}; }
    FR_HD Cplx operator-(Cplx a, double b) { return {a.re - b, a.im}; }
    FR_HD Cplx operator-(double a, Cplx b) { return {a - b.re, -b.im}; }
    FR_HD Cplx operator*(Cplx a, double b) { return {a.re * b, a.im * b}; }
    FR_HD Cplx operator*(double a, Cplx b) { return {a * b.re, a * b.im}; }
    FR_HD Cplx operator/(Cplx a, double b) { return {a.re / b, a.im / b}; }
    FR_HD Cplx operator/(double a, Cplx b) { return Cplx{a, 0.0} / b; }
    FR_HD Cplx sqrt(Cplx z)
    {
        // principal square root, cancellation-free in both half planes
        const double m = ::sqrt(z.re * z.re + z.im * z.im);
        if (z.re >= 0.0)
        {
            const double t = ::sqrt(0.5 * (m + z.re));
            return {t, (t == 0.0) ? 0.0 : z.im / (2.0 * t)};
        }
        const double t = ::sqrt(0.5 * (m - z.re));
        return {::fabs(z.im) / (2.0 * t), ::copysign(t, z.im)};
    }

    // ------------- uniform helpers over the three types -------------
    template <class S>
    struct Sc;
    template <>
    struct Sc<double>
    {
        static FR_HD double c(double x) { return x; }
        static FR_HD double branch(double x) { return x; }
        static FR_HD double val(double x) { return x; }
        static FR_HD double wrap_positive(double x) { return (DBL_MIN < x) ? x : DBL_MIN; }
        static FR_HD double frozen(double x) { return x; }
        static FR_HD double sel(bool k, double a, double b) { return k ? a : b; }
    };
    template <>
    struct Sc<Dual>
    {
        static FR_HD Dual c(double x) { return {x, 0.0}; }
        static FR_HD double branch(Dual x) { return x.v; }
        static FR_HD double val(Dual x) { return x.v; }
        static FR_HD Dual wrap_positive(Dual x)
        {
            const bool k = DBL_MIN < x.v; // component-wise selects keep aggregates out of memory
            return {k ? x.v : DBL_MIN, k ? x.d : 0.0};
        }
        static FR_HD Dual frozen(Dual x) { return {x.v, 0.0}; }
        static FR_HD Dual sel(bool k, Dual a, Dual b) { return {k ? a.v : b.v, k ? a.d : b.d}; }
    };
    template <>
    struct Sc<Cplx>
    {
        static FR_HD Cplx c(double x) { return {x, 0.0}; }
        static FR_HD double branch(Cplx x) { return x.re + x.im; }
        static FR_HD double val(Cplx x) { return x.re; }
        static FR_HD Cplx wrap_positive(Cplx x)
        {
            const bool k = x.re < DBL_MIN;
            return {k ? DBL_MIN : x.re, k ? 0.0 : x.im};
        }
        static FR_HD Cplx frozen(Cplx x) { return {x.re, 0.0}; }
        static FR_HD Cplx sel(bool k, Cplx a, Cplx b) { return {k ? a.re : b.re, k ? a.im : b.im}; }
    };
    // Dual: the value part as for doubles (the same primal as the residual)
    FR_HD Dual rcp_(Dual x)
    {
        const double r = rcp_(x.v);
        return {r, -(r * r) * x.d};
    }
    FR_HD Dual quot_(Dual a, Dual b)
    {
        const double r = rcp_(b.v);
        const double q = quot_(a.v, b.v);
        return {q, (a.d - q * b.d) * r};
    }
    FR_HD Cplx rcp_(Cplx x) { return 1.0 / x; }
    FR_HD Cplx quot_(Cplx a, Cplx b) { return a / b; }
    FR_HD Dual sqrt_(Dual x) { return sqrt(x); }
    FR_HD Cplx sqrt_(Cplx x) { return sqrt(x); }
    // exponential (change of variables of the log-variable implicit path)
    FR_HD double exp_(double x) { return ::exp(x); }
    FR_HD Dual exp_(Dual x)
    {
        const double e = ::exp(x.v);
        return {e, x.d * e};
    }
    FR_HD Cplx exp_(Cplx z)
    {
        const double m = ::exp(z.re);
        return {m * ::cos(z.im), m * ::sin(z.im)};
    }
} // namespace frsdg

// Normalized Planck and Rosseland integrals (host and device), restating
// src/planck.hpp:1-253 from the published method it cites (B. A. Clark,
// "Computing multigroup radiation integrals using polylogarithm-based
// methods", JCP 1987): the normalized Planck integral
//   P(x) = (15 / pi^4) int_0^x t^3 / (e^t - 1) dt
// is the smaller of its Taylor series about 0 (terms through x^21, whose
// coefficients are B_n / (n! (n + 3)) with the Bernoulli numbers B_n) and
// the polylogarithm form 1 - (15 / pi^4) (6 Li4 + 6 x Li3 + 3 x^2 Li2 +
// x^3 Li1)(e^-x) with each Li_a summed over its first 9 terms.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace allium {

__host__ __device__ inline double planck_norm() { return 15.0 / (M_PI * M_PI * M_PI * M_PI); }

// Taylor series of int_0^x t^3/(e^t - 1) dt times 15/pi^4
__host__ __device__ inline double planck_taylor(double x)
{
    // B_n / (n! (n + 3)) for n = 0, 1, 2, 4, ..., 18
    const double c3 = 1.0 / 3.0, c4 = -1.0 / 8.0, c5 = 1.0 / 60.0;
    const double c7 = -1.0 / 5040.0;                                 // B4 = -1/30
    const double c9 = 1.0 / 272160.0;                                // B6 = 1/42
    const double c11 = -1.0 / 13305600.0;                            // B8 = -1/30
    const double c13 = 1.0 / 622702080.0;                            // B10 = 5/66
    const double c15 = -691.0 / (2730.0 * 479001600.0 * 15.0);       // B12 = -691/2730
    const double c17 = 7.0 / (6.0 * 87178291200.0 * 17.0);           // B14 = 7/6
    const double c19 = -3617.0 / (510.0 * 20922789888000.0 * 19.0);  // B16 = -3617/510
    const double c21 = 43867.0 / (798.0 * 6402373705728000.0 * 21.0);  // B18 = 43867/798
    const double x2 = x * x;
    double t = x2 * c21 + c19;
    t = t * x2 + c17;
    t = t * x2 + c15;
    t = t * x2 + c13;
    t = t * x2 + c11;
    t = t * x2 + c9;
    t = t * x2 + c7;
    t = t * x2 + c5;
    t = t * x + c4;
    t = t * x + c3;
    return t * x * x2 * planck_norm();
}

__host__ __device__ inline double planck_polylog(double x)
{
    const double e = exp(-x);
    double ep = e, li1 = e, li2 = e, li3 = e, li4 = e;
    for (int l = 2; l <= 9; ++l) {
        const double inv = 1.0 / l;
        ep *= e;
        double t = ep * inv;
        li1 += t;
        t *= inv;
        li2 += t;
        t *= inv;
        li3 += t;
        t *= inv;
        li4 += t;
    }
    return 1.0 - planck_norm() * (6.0 * li4 + x * (6.0 * li3 + x * (3.0 * li2 + x * li1)));
}

// normalized Planck integral over [0, x]
__host__ __device__ inline double planck_integral(double x)
{
    if (x > 1e100) return 1.0;
    if (x == 0.0) return 0.0;
    return fmin(planck_taylor(x), planck_polylog(x));
}

// normalized Rosseland integral over [0, x]: P(x) - (15/pi^4)/4 x^4/(e^x - 1)
// (its limit 0 at x = 0, where the closed form reads 0/0)
__host__ __device__ inline double rosseland_integral(double x)
{
    if (x == 0.0) return 0.0;
    const double p = planck_integral(x);
    if (x > 1.1579208923731620e77) return p;   // (DBL_MAX)^(1/4)
    return p - 0.25 * planck_norm() * (x * x) * (x * x) / expm1(x);
}

// fraction of the Planck (R = 0) or Rosseland (R = 1) spectrum at T in each
// of G groups; the first group starts at 0 and the last extends to infinity
template <int R>
__host__ __device__ inline void planck_spectrum(const double *bounds, int G, double T, double *out)
{
    double prev = 0.0;
    for (int g = 0; g < G - 1; ++g) {
        const double x = bounds[g + 1] / T;
        const double next = R ? rosseland_integral(x) : planck_integral(x);
        out[g] = next - prev;
        prev = next;
    }
    out[G - 1] = 1.0 - prev;
}

// the same fractions handed to f(g, fraction) in group order, without an
// array (a per-thread array indexed by a runtime group count lives in scratch)
template <int R, typename F>
__host__ __device__ inline void planck_spectrum_each(const double *bounds, int G, double T, F &&f)
{
    double prev = 0.0;
    for (int g = 0; g < G - 1; ++g) {
        const double x = bounds[g + 1] / T;
        const double next = R ? rosseland_integral(x) : planck_integral(x);
        f(g, next - prev);
        prev = next;
    }
    f(G - 1, 1.0 - prev);
}

}  // namespace allium

// Regularised delta kernels with 4-point support, evaluated on the four
// stencil offsets r = d + s, s in {-1, 0, 1, 2}, d in (-1, 0] (the stencil
// of ib/delta.h:113-147 for meshwidths = 4).
#pragma once
#include "geometry.hpp"
#include <math.h>

namespace ibm {

// Cosine kernel phi(r) = (1 + cos(pi r / 2)) / 4 (ib/cosine.h:9-16).  The four
// values need one sine and one cosine: with t = pi d / 2,
//   phi(d-1) = (1 + sin t)/4, phi(d) = (1 + cos t)/4,
//   phi(d+1) = (1 - sin t)/4, phi(d+2) = (1 - cos t)/4.
// t lies in (-pi/2, 0], so no argument reduction is needed: the Taylor series
// of sin and cos at the half angle u = t/2 (|u| <= pi/4, truncated after
// u^17/17! and u^16/16!, remainder below 1e-19) and the double-angle
// identities sin t = 2 sin u cos u, cos t = 1 - 2 sin^2 u give both values
// in about twenty fused multiply-adds.
IBM_HD void half_angle_sincos(double t, double* s_out, double* c_out)
{
    double u = 0.5 * t;
    double v = u * u;
    // sp = -(sin u)/u = sum_k (-1)^(k+1) v^k / (2k+1)!, highest term first
    double sp = -1.0 / 355687428096000.0;                  // 17!
    sp = fma(sp, v, 1.0 / 1307674368000.0);                 // 15!
    sp = fma(sp, v, -1.0 / 6227020800.0);                   // 13!
    sp = fma(sp, v, 1.0 / 39916800.0);                      // 11!
    sp = fma(sp, v, -1.0 / 362880.0);                       // 9!
    sp = fma(sp, v, 1.0 / 5040.0);                          // 7!
    sp = fma(sp, v, -1.0 / 120.0);                          // 5!
    sp = fma(sp, v, 1.0 / 6.0);                             // 3!
    sp = fma(sp, v, -1.0);
    double su = -u * sp;
    // cos u = sum_k (-1)^k v^k / (2k)!
    double cp = 1.0 / 20922789888000.0;                     // 16!
    cp = fma(cp, v, -1.0 / 87178291200.0);                  // 14!
    cp = fma(cp, v, 1.0 / 479001600.0);                     // 12!
    cp = fma(cp, v, -1.0 / 3628800.0);                      // 10!
    cp = fma(cp, v, 1.0 / 40320.0);                         // 8!
    cp = fma(cp, v, -1.0 / 720.0);                          // 6!
    cp = fma(cp, v, 1.0 / 24.0);                            // 4!
    cp = fma(cp, v, -0.5);                                  // 2!
    double cu = fma(cp, v, 1.0);
    *s_out = 2.0 * su * cu;
    *c_out = fma(-2.0 * su, su, 1.0);
}

IBM_HD void delta_cosine4(double d, double w[4])
{
    double s, c;
    half_angle_sincos(M_PI_2 * d, &s, &c);
    w[0] = 0.25 * (1 + s);
    w[1] = 0.25 * (1 + c);
    w[2] = 0.25 * (1 - s);
    w[3] = 0.25 * (1 - c);
}

// Peskin's 4-point kernel (ib/peskin.h:6-14), evaluated as written.
IBM_HD double delta_peskin(double r)
{
#pragma clang fp contract(off)
    double s = r < 0 ? -r : r;
    return s < 1 ? (3 - 2 * s + sqrt(1 + 4 * s - 4 * s * s)) / 8
                 : (5 - 2 * s - sqrt(-7 + 12 * s - 4 * s * s)) / 8;
}

// Cardinal B-spline by the reference's recursion (ib/bspline.h:13-27):
//   B_0(r) = [-1/2 <= r < 1/2],
//   B_n(r) = ((n+1)/(2n) + r/n) B_{n-1}(r + 1/2) + ((n+1)/(2n) - r/n) B_{n-1}(r - 1/2).
template <int n>
IBM_HD double delta_bspline(double r)
{
#pragma clang fp contract(off)
    if constexpr (n == 0) {
        return (-0.5 <= r && r < 0.5) ? 1.0 : 0.0;
    } else {
        constexpr double b = (n + 1.0) / (2.0 * n);
        double a = r / n;
        return (b + a) * delta_bspline<n - 1>(r + 0.5) + (b - a) * delta_bspline<n - 1>(r - 0.5);
    }
}

IBM_HD void delta4(int kind, double d, double w[4])
{
    if (kind == IBM_DELTA_COSINE) {
        delta_cosine4(d, w);
    } else if (kind == IBM_DELTA_PESKIN) {
#pragma unroll
        for (int s = 0; s < 4; ++s) w[s] = delta_peskin(d + (s - 1));
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) w[s] = delta_bspline<3>(d + (s - 1));
    }
}

} // namespace ibm

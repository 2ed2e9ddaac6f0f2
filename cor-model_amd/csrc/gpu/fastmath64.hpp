// fastmath64.hpp -- FP64 e^x and sin/cos for the mixed tracer's cancelling harmonic
// sums (the TS05 shield fields and warped-tail sums, whose terms are 1e4-1e5 times
// their sum and so stay FP64 in the mixed path).
//
// Cody-Waite argument reduction with a two-part constant, then Taylor polynomials in
// Horner form (J.-M. Muller, Elementary Functions, ch. 6 and 11):
//   e^x = 2^n e^r,            n = rint(x / ln 2),    |r| <= ln(2)/2, series to r^13;
//   sin, cos x from r,        k = rint(x / (pi/2)),  |r| <= pi/4,    series to r^15 / r^16,
//                             and the quadrant k mod 4.
// Truncation is below 1e-16 relative; each result is within a few ulp.  One reduction
// serves both sin and cos.  The reduction is accurate for |x| up to ~1e5 (these sums see
// |x| < 1e3); NaN propagates.  Host and device compile the same arithmetic (fma, rint,
// ldexp), so the CPU tests check exactly what the GPU runs.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define COR_FM_HD __host__ __device__
#else
#define COR_FM_HD
#endif

namespace cor {
namespace fm64 {

// ln 2 and pi/2, each as the nearest double plus the nearest double to the remainder
constexpr double LN2_HI = 0.6931471805599453;
constexpr double LN2_LO = 2.3190468138462996e-17;
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double PIO2_HI = 1.5707963267948966;
constexpr double PIO2_LO = 6.123233995736766e-17;
constexpr double INV_PIO2 = 0.6366197723675814;

COR_FM_HD inline double exp(double x) {
    // beyond [-746, 710] the result is 0 or +inf anyway; the clamp keeps n an int
    const double xc = x < -746.0 ? -746.0 : (x > 710.0 ? 710.0 : x);
    const double n = rint(xc * INV_LN2);
    double r = fma(-n, LN2_HI, xc);
    r = fma(-n, LN2_LO, r);
    double p = 1.0 / 6227020800.0;          // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);       // 1/12!
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, n == n ? (int)n : 0);   // a NaN x leaves p NaN
}

COR_FM_HD inline void sincos(double x, double *s, double *c) {
    const double k = rint(x * INV_PIO2);
    double r = fma(-k, PIO2_HI, x);
    r = fma(-k, PIO2_LO, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! - ...)), to r^15
    double ps = -1.0 / 1307674368000.0;     // -1/15!
    ps = fma(ps, r2, 1.0 / 6227020800.0);
    ps = fma(ps, r2, -1.0 / 39916800.0);
    ps = fma(ps, r2, 1.0 / 362880.0);
    ps = fma(ps, r2, -1.0 / 5040.0);
    ps = fma(ps, r2, 1.0 / 120.0);
    ps = fma(ps, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, ps, r);
    // cos r = 1 + r^2 (-1/2 + r^2 (1/4! - ...)), to r^16
    double pc = 1.0 / 20922789888000.0;     // 1/16!
    pc = fma(pc, r2, -1.0 / 87178291200.0);
    pc = fma(pc, r2, 1.0 / 479001600.0);
    pc = fma(pc, r2, -1.0 / 3628800.0);
    pc = fma(pc, r2, 1.0 / 40320.0);
    pc = fma(pc, r2, -1.0 / 720.0);
    pc = fma(pc, r2, 1.0 / 24.0);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    // x = r + k pi/2: quadrant k mod 4 (as a double, so a NaN k needs no conversion)
    const double q = k - 4.0 * floor(k * 0.25);
    if (q == 0.0) {
        *s = sr;
        *c = cr;
    } else if (q == 1.0) {
        *s = cr;
        *c = -sr;
    } else if (q == 2.0) {
        *s = -sr;
        *c = -cr;
    } else {
        *s = -cr;
        *c = sr;
    }
}

}  // namespace fm64
}  // namespace cor

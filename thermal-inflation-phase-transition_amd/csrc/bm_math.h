// bm_math.h -- float64 log, sqrt and cos / sin for the hash Box-Muller draw, on the draw's own
// domains (u1 in [1e-12, 1), -2 log u1 in (0, 56), the angle 2 pi u2 with u2 in [0, 1)).
//
// The generic device libm routines handle every argument (cos of any magnitude needs a
// Payne-Hanek reduction, sqrt and division scale for denormals and overflow); here the
// arguments are bounded, so each function is a short, branch-free sequence:
//
//  * -2 log(x), x a positive normal number: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
//    s = (m - 1) / (m + 1), log m = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ...) (the atanh
//    series; |s| <= 0.1716, the terms to s^22 leave < 1e-18 relative), and
//    log x = e ln2_hi + (e ln2_lo + log m) with ln 2 split so that e ln2_hi is exact.
//    The quotient is f * y with y the hardware reciprocal of m + 1 after one Newton step,
//    then one residual correction q + (f - d q) y, which leaves the square of y's error
//    (Markstein's division; d in [1.7, 2.5] needs none of the generic routine's scaling).
//  * sqrt(t), 0 < t < 2^64 (normal): y = hardware 1/sqrt(t), one coupled Newton step for
//    g ~ sqrt(t) and h ~ 1/(2 sqrt(t)) (Goldschmidt; it squares the relative error), then one
//    residual correction g + (t - g^2) h.
//  * cos(2 pi u) for u in [0, 1): with v = min(u, 1 - u) (exact for a u of 53 bits),
//    cos(2 pi u) = 1 - 2 sin^2(pi v) and pi v in [0, pi/2], sin by its Taylor series to the
//    x^21 term (the next term is below 1.2e-18 there).  For v = u, pi v is exactly half of
//    the rounded 2 pi u (the two constants are one binary exponent apart).
//  * cos(x), sin(x) together for 0 <= x < 2^20 (the complex field's pair): k = nearest integer
//    to x / (pi/2), r = x - k pi/2 by Cody and Waite's two-constant reduction
//    (pi/2 = pio2_hi + pio2_lo, |r| <= pi/4), the Taylor series of sin r (to r^21) and cos r
//    (to r^22) on |r| <= pi/4, and the quadrant k mod 4 selecting +-sin r / +-cos r.
//
// Error: a few ulp for log and sqrt over their domains and < 6e-16 absolute for cos / sin;
// the draw z = sqrt(-2 log u1) cos(2 pi u2) then agrees with the host libm's to ~1e-15 of
// sqrt(-2 log u1).  Products and sums are written with explicit fma so the result does not
// depend on floating-point contraction flags.
#pragma once

#ifndef BM_HD
#define BM_HD __device__
#endif

namespace tip {

// -2 log(x) for a positive normal x (the Box-Muller radius squared); the factor -2 is carried
// in the constants, which changes no rounding (powers of two)
BM_HD inline double bm_neg2log(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);  // x = m 2^e, m in [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m * 2.0 : m;     // exact
    e = lo ? e - 1 : e;
    const double f = m - 1.0;  // exact (Sterbenz)
    const double d = m + 1.0;
    double y = __builtin_amdgcn_rcp(d);
    y = fma(y, fma(-d, y, 1.0), y);
    const double q = f * y;
    const double s = fma(fma(-d, q, f), y, q);
    const double s2 = s * s;
    // 1 + s^2/3 + s^4/5 + ... + s^20/21, Horner in s^2
    double p = 1.0 / 21.0;
    p = fma(p, s2, 1.0 / 19.0);
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    const double m4s = -4.0 * s;
    const double neg2logm = fma(m4s * s2, p, m4s);  // -2 (2 s + 2 s^3 p)
    // ln 2 = ln2_hi + ln2_lo, ln2_hi with its low 21 bits zero: e * ln2_hi is exact for
    // |e| < 2^21
    const double m2ln2_hi = -2.0 * 6.93147180369123816490e-01;
    const double m2ln2_lo = -2.0 * 1.90821492927058770002e-10;
    const double de = (double)e;
    return fma(de, m2ln2_hi, fma(de, m2ln2_lo, neg2logm));
}

// sqrt(t) for a positive normal t below 2^64
BM_HD inline double bm_sqrt(double t)
{
    const double y = __builtin_amdgcn_rsq(t);
    double g = t * y;
    double h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    return fma(fma(-g, g, t), h, g);
}

// cos(2 pi u) for u in [0, 1)
BM_HD inline double bm_cos2pi(double u)
{
    const double v = u <= 0.5 ? u : 1.0 - u;
    const double x = 3.141592653589793 * v;  // pi v, in [0, pi/2]
    const double x2 = x * x;
    // sin x = x - x^3 p, p = 1/3! - x^2/5! + x^4/7! - ... - x^18/21!, Horner in x^2
    double p = -1.0 / 51090942171709440000.0;   // -1/21!
    p = fma(p, x2, 1.0 / 121645100408832000.0);  // 1/19!
    p = fma(p, x2, -1.0 / 355687428096000.0);    // -1/17!
    p = fma(p, x2, 1.0 / 1307674368000.0);       // 1/15!
    p = fma(p, x2, -1.0 / 6227020800.0);         // -1/13!
    p = fma(p, x2, 1.0 / 39916800.0);            // 1/11!
    p = fma(p, x2, -1.0 / 362880.0);             // -1/9!
    p = fma(p, x2, 1.0 / 5040.0);                // 1/7!
    p = fma(p, x2, -1.0 / 120.0);                // -1/5!
    p = fma(p, x2, 1.0 / 6.0);                   // 1/3!, subtracted below
    const double sn = fma(-(x * x2), p, x);
    return fma(-2.0 * sn, sn, 1.0);
}

// cos and sin of x in [0, 2^20)
BM_HD inline void bm_sincos(double x, double* sn, double* cs)
{
    const double two_over_pi = 6.36619772367581382433e-01;
    const double pio2_hi = 1.57079632679489655800e+00;   // nearest double to pi/2
    const double pio2_lo = 6.12323399573676603587e-17;   // pi/2 - pio2_hi
    const double k = rint(x * two_over_pi);
    double r = fma(-k, pio2_hi, x);
    r = fma(-k, pio2_lo, r);
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ... + r^20/21!)
    double ps = 1.0 / 51090942171709440000.0;   // 1/21!
    ps = fma(ps, r2, -1.0 / 121645100408832000.0);  // -1/19!
    ps = fma(ps, r2, 1.0 / 355687428096000.0);      // 1/17!
    ps = fma(ps, r2, -1.0 / 1307674368000.0);       // -1/15!
    ps = fma(ps, r2, 1.0 / 6227020800.0);           // 1/13!
    ps = fma(ps, r2, -1.0 / 39916800.0);            // -1/11!
    ps = fma(ps, r2, 1.0 / 362880.0);               // 1/9!
    ps = fma(ps, r2, -1.0 / 5040.0);                // -1/7!
    ps = fma(ps, r2, 1.0 / 120.0);                  // 1/5!
    ps = fma(ps, r2, -1.0 / 6.0);                   // -1/3!
    const double sr = fma(r * r2, ps, r);
    // cos r = 1 - r^2/2! + r^4/4! - ... + r^22/22!
    double pc = 1.0 / 1124000727777607680000.0;     // 1/22!
    pc = fma(pc, r2, -1.0 / 2432902008176640000.0); // -1/20!
    pc = fma(pc, r2, 1.0 / 6402373705728000.0);     // 1/18!
    pc = fma(pc, r2, -1.0 / 20922789888000.0);      // -1/16!
    pc = fma(pc, r2, 1.0 / 87178291200.0);          // 1/14!
    pc = fma(pc, r2, -1.0 / 479001600.0);           // -1/12!
    pc = fma(pc, r2, 1.0 / 3628800.0);              // 1/10!
    pc = fma(pc, r2, -1.0 / 40320.0);               // -1/8!
    pc = fma(pc, r2, 1.0 / 720.0);                  // 1/6!
    pc = fma(pc, r2, -1.0 / 24.0);                  // -1/4!
    pc = fma(pc, r2, 0.5);                          // the 1/2! term, subtracted below
    const double cr = fma(-r2, pc, 1.0);
    const int q = (int)k & 3;
    // cos(x) = cos(r + q pi/2), sin(x) = sin(r + q pi/2)
    const double c0 = (q & 1) ? sr : cr;
    const double s0 = (q & 1) ? cr : sr;
    *cs = ((q + 1) & 2) ? -c0 : c0;
    *sn = (q & 2) ? -s0 : s0;
}

}  // namespace tip

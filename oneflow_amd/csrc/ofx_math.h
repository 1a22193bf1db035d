// ofx_math.h -- x^y for positive x with the same bits on the host and the
// device.  The library's kernels and the CPU oracle both call ofx_pow where
// the reference calls pow with a non-integer exponent (the SA wall function
// fw, codes/turb/src/TurbCom.cpp:817, and the farfield characteristic
// state, codes/ns/src/NsBcSolver.cpp:237-245): every operation below is an IEEE
// double operation or an explicit fma, correctly rounded on both sides and
// evaluated in the written order (both builds use -ffp-contract=off), so the
// two agree bit for bit, where a host libm and the device math library may
// differ in the last place.
//
// Method: ln x in double-double (x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), the atanh series in s^2 <= 0.0295
// to the 25th power), z = y ln x in double-double, then e^z with the
// Cody-Waite reduction z = k ln 2 + r (ln 2 split into a 32-bit head and a
// tail, both derived from ln 2 to 60 digits) and the Taylor series of e^r on
// |r| <= ln 2 / 2 to r^13, its terms up to r^2 in double-double.  The error is within one unit in the last place
// for the arguments the solver uses (checked against the host libm over wide
// ranges); it equals libm's correctly rounded pow in most cases.
//
// Include with OFX_HD defined as the function qualifiers of the build
// (device code: host and device, force-inlined); without it the functions are
// plain inline host functions.
#pragma once
#include <math.h>

#ifndef OFX_HD
#define OFX_HD inline
#endif

namespace ofxm {

struct dd {
    double hi, lo;
};

OFX_HD dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}
// requires |a| >= |b| (or a == 0)
OFX_HD dd fast_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    return dd{s, e};
}
OFX_HD dd two_prod(double a, double b) {
    double p = a * b;
    return dd{p, fma(a, b, -p)};
}

// ln 2 to double-double, and split into a 32-bit head (k * head exact for
// |k| < 2^21) and a tail
constexpr double LN2_HI = 0.6931471805599453;
constexpr double LN2_LO = 2.3190468138462996e-17;
constexpr double LN2_H32 = 0.6931471806019545;
constexpr double LN2_T32 = -4.2009150726810846e-11;
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double SQRT_HALF = 0.70710678118654752;

// ln x for finite normal or subnormal x > 0
OFX_HD dd log_dd(double x) {
    int e;
    double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
    if (m < SQRT_HALF) {
        m = m * 2.0;
        e = e - 1;
    }
    double f = m - 1.0;            // exact: m in [sqrt(1/2), sqrt(2))
    dd d = two_sum(m, 1.0);        // m + 1 exactly
    double s = f / d.hi;
    double r = fma(-s, d.hi, f) - s * d.lo;  // f - s (m + 1)
    double sl = r / d.hi;                    // s + sl = f / (m + 1)
    double t = s * s;
    // sum_{k>=1} t^(k-1) / (2k + 1), k = 1..12
    double p = 1.0 / 25.0;
    p = p * t + 1.0 / 23.0;
    p = p * t + 1.0 / 21.0;
    p = p * t + 1.0 / 19.0;
    p = p * t + 1.0 / 17.0;
    p = p * t + 1.0 / 15.0;
    p = p * t + 1.0 / 13.0;
    p = p * t + 1.0 / 11.0;
    p = p * t + 1.0 / 9.0;
    p = p * t + 1.0 / 7.0;
    p = p * t + 1.0 / 5.0;
    p = p * t + 1.0 / 3.0;
    double tail = 2.0 * s * t * p;
    dd lm = fast_two_sum(2.0 * s, 2.0 * sl + tail);  // ln m
    dd el = two_prod((double)e, LN2_HI);             // exact
    double el_lo = el.lo + (double)e * LN2_LO;
    dd s1 = two_sum(el.hi, lm.hi);
    double lo = s1.lo + (el_lo + lm.lo);
    return fast_two_sum(s1.hi, lo);
}

// e^(zh + zl) for |zh| below the overflow threshold
OFX_HD double exp_dd(double zh, double zl) {
    double k = floor(zh * INV_LN2 + 0.5);
    dd rr = two_sum(zh, -(k * LN2_H32));  // k * LN2_H32 exact
    double rlo = rr.lo + (zl - k * LN2_T32);
    dd rd = two_sum(rr.hi, rlo);          // r = rd.hi + rd.lo, |r| <= ln 2 / 2 (+ tiny)
    double x = rd.hi;
    // e^r = 1 + x + x^2/2 + x^3 (1/3! + x/4! + ... + x^10/13!) + rd.lo (1 + x):
    // the leading terms in double-double, the cubic tail (< 0.007) in double
    double q = 1.0 / 6227020800.0;
    q = q * x + 1.0 / 479001600.0;
    q = q * x + 1.0 / 39916800.0;
    q = q * x + 1.0 / 3628800.0;
    q = q * x + 1.0 / 362880.0;
    q = q * x + 1.0 / 40320.0;
    q = q * x + 1.0 / 5040.0;
    q = q * x + 1.0 / 720.0;
    q = q * x + 1.0 / 120.0;
    q = q * x + 1.0 / 24.0;
    q = q * x + 1.0 / 6.0;
    dd x2 = two_prod(x, x);
    double q3 = q * (x2.hi * x);
    dd s1 = fast_two_sum(x, 0.5 * x2.hi);  // |x| >= x^2 / 2 for |x| <= ln 2 / 2
    double lo = s1.lo + ((0.5 * x2.lo + q3) + (rd.lo + rd.lo * x));
    dd a = fast_two_sum(1.0, s1.hi);
    return ldexp(a.hi + (a.lo + lo), (int)k);
}

// x^y for x > 0 finite; other x (zero, negative, infinite, NaN) follow the
// host / device pow, whose special values agree.
OFX_HD double ofx_pow(double x, double y) {
    if (!(x > 0.0) || !(x < INFINITY) || !(y == y)) return pow(x, y);
    if (x == 1.0) return 1.0;
    dd L = log_dd(x);
    dd z = two_prod(y, L.hi);
    double zl = z.lo + y * L.lo;
    dd zz = fast_two_sum(z.hi, zl);
    if (zz.hi > 709.0 || zz.hi < -708.0) return pow(x, y);  // overflow / subnormal range
    return exp_dd(zz.hi, zz.lo);
}

}  // namespace ofxm

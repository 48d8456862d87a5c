// Single-precision power and logarithm for the routing physics, evaluated in
// double precision to about 2^-45 relative error and rounded once to fp32.
//
// The reference CPU path calls libm powf/logf, whose results are correctly
// rounded on all but a small fraction of inputs; an fp32 result from a ~2^-45
// accurate double value is correctly rounded except within 2^-21 ulp of a
// rounding midpoint, so these agree with libm as often as libm is correctly
// rounded. They cost about a third of a full-precision double pow.
//
// Methods (textbook): ln x from x = m 2^e with m in [1/sqrt2, sqrt2),
// ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, as an odd series to s^19;
// e^t by Cody-Waite reduction t = k ln2 + r, |r| <= ln2/2, and the Taylor
// polynomial of e^r to degree 12. Zero, negative, infinite and NaN operands
// take the C99 special-case results without a library call.
// Every multiply-add is an explicit fma, so the host and the device produce
// the same bits. pow_f32 and log_f32 stay out of line: inlined at every call
// site of the adjoint they raised its register use from 128 to 169 VGPRs.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define DDRS_HD __host__ __device__
#else
#define DDRS_HD
#endif

// pow_f32 / log_f32 / pow_ln_f32 are out of line (a call makes the caller
// wait for its outstanding memory loads, but inlined at every call site the
// powers raise register use); pow_f32_inl / pow_ln_f32_inl are the same
// evaluations for callers with registers to spare, where the polynomial
// constants stay in registers across calls.
#define DDRS_MATH_LINKAGE __attribute__((noinline))

namespace ddrs {

// A wave-uniform double constant placed in a scalar register pair: the vector
// multiply-adds then take it as an operand, instead of each constant being
// rebuilt in vector registers with two vector moves (same values, same bits).
DDRS_HD inline double kc(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+s"(c));
#endif
    return c;
}

// ln(x) for finite x > 0 (double), ~2^-50 relative (absolute near x = 1)
DDRS_HD inline double ln_reduced(double x) {
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < kc(0.70710678118654752440)) {
        m = m + m;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = kc(1.0 / 19.0);
    p = fma(p, z, kc(1.0 / 17.0));
    p = fma(p, z, kc(1.0 / 15.0));
    p = fma(p, z, kc(1.0 / 13.0));
    p = fma(p, z, kc(1.0 / 11.0));
    p = fma(p, z, kc(1.0 / 9.0));
    p = fma(p, z, kc(1.0 / 7.0));
    p = fma(p, z, kc(1.0 / 5.0));
    p = fma(p, z, kc(1.0 / 3.0));
    const double lnm = fma(2.0 * s * z, p, 2.0 * s);  // 2 s (1 + z p)
    const double ln2_hi = kc(6.93147180369123816490e-01);
    const double ln2_lo = kc(1.90821492927058770002e-10);
    const double de = static_cast<double>(e);
    return fma(de, ln2_hi, fma(de, ln2_lo, lnm));
}

// e^t for |t| <= 800 (overflow to inf and underflow to 0 through ldexp), ~2^-50 relative
DDRS_HD inline double exp_reduced(double t) {
    const double inv_ln2 = kc(1.44269504088896338700e+00);
    const double ln2_hi = kc(6.93147180369123816490e-01);
    const double ln2_lo = kc(1.90821492927058770002e-10);
    const double k = rint(t * inv_ln2);
    const double r = fma(-k, ln2_lo, fma(-k, ln2_hi, t));
    double p = kc(1.0 / 479001600.0);    // 1/12!
    p = fma(p, r, kc(1.0 / 39916800.0)); // 1/11!
    p = fma(p, r, kc(1.0 / 3628800.0));
    p = fma(p, r, kc(1.0 / 362880.0));
    p = fma(p, r, kc(1.0 / 40320.0));
    p = fma(p, r, kc(1.0 / 5040.0));
    p = fma(p, r, kc(1.0 / 720.0));
    p = fma(p, r, kc(1.0 / 120.0));
    p = fma(p, r, kc(1.0 / 24.0));
    p = fma(p, r, kc(1.0 / 6.0));
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, static_cast<int>(k));
}

// C99 pow results for the operands the fast path leaves out (x <= 0, x = +inf,
// NaN or infinite y), for the cases the physics can produce: x = 0, x = +inf,
// NaN, negative x with y = 2 (exact square) and negative x otherwise (NaN).
DDRS_HD inline float pow_special(double xd, double yd) {
    if (xd != xd || yd != yd) return (yd == 0.0) ? 1.0f : static_cast<float>(xd + yd);
    if (yd == 0.0) return 1.0f;
    if (xd == 0.0) return (yd > 0.0) ? 0.0f : static_cast<float>(INFINITY);
    if (xd == INFINITY) return (yd > 0.0) ? static_cast<float>(INFINITY) : 0.0f;
    if (yd == 2.0) return static_cast<float>(xd * xd);
    return static_cast<float>(NAN);
}

// C99 log for x <= 0, x = +inf and NaN
DDRS_HD inline float log_special(double xd) {
    if (xd == 0.0) return -static_cast<float>(INFINITY);
    if (xd == INFINITY) return static_cast<float>(INFINITY);
    return static_cast<float>(NAN);
}

DDRS_HD inline bool pow_fast_path(double xd, double yd) {
    return xd > 0.0 && xd < INFINITY && yd == yd && fabs(yd) < INFINITY;
}

// e^(y ln x) rounded once to fp32
DDRS_HD inline float pow_from_ln(double l, double yd) {
    double t = yd * l;
    t = fmin(fmax(t, -800.0), 800.0);  // beyond this the fp32 result is 0 or inf
    return static_cast<float>(exp_reduced(t));
}

// x^y for fp32 x, y, rounded once to fp32.
DDRS_HD inline __attribute__((always_inline)) float pow_f32_inl(float x, float y) {
    const double xd = static_cast<double>(x), yd = static_cast<double>(y);
    if (pow_fast_path(xd, yd)) return pow_from_ln(ln_reduced(xd), yd);
    return pow_special(xd, yd);
}
DDRS_HD inline DDRS_MATH_LINKAGE float pow_f32(float x, float y) { return pow_f32_inl(x, y); }

// ln x for fp32 x, rounded once to fp32
DDRS_HD inline DDRS_MATH_LINKAGE float log_f32(float x) {
    const double xd = static_cast<double>(x);
    if (xd > 0.0 && xd < INFINITY) return static_cast<float>(ln_reduced(xd));
    return log_special(xd);
}

// pow_f32(x, y) and log_f32(x) from one evaluation of ln x; both results are
// bit-identical to the separate calls.
struct PowLn {
    float pw, ln;
};
DDRS_HD inline __attribute__((always_inline)) PowLn pow_ln_f32_inl(float x, float y) {
    const double xd = static_cast<double>(x), yd = static_cast<double>(y);
    PowLn r;
    if (xd > 0.0 && xd < INFINITY) {
        const double l = ln_reduced(xd);
        r.ln = static_cast<float>(l);
        r.pw = pow_fast_path(xd, yd) ? pow_from_ln(l, yd) : pow_special(xd, yd);
    } else {
        r.ln = log_special(xd);
        r.pw = pow_special(xd, yd);
    }
    return r;
}
DDRS_HD inline DDRS_MATH_LINKAGE PowLn pow_ln_f32(float x, float y) { return pow_ln_f32_inl(x, y); }

}  // namespace ddrs

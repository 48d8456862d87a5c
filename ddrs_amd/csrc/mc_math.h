// Single-precision power and logarithm for the routing physics, evaluated in
// double precision to about 2^-50 relative error and rounded once to fp32.
//
// The reference CPU path calls libm powf/logf, whose results are correctly
// rounded on all but a small fraction of inputs; an fp32 result from a ~2^-50
// accurate double value is correctly rounded except within 2^-26 ulp of a
// rounding midpoint, so these agree with libm as often as libm is correctly
// rounded. They cost about a third of a full-precision double pow.
//
// Methods (textbook): ln x by table reduction, x = 2^k z with z in
// [0.7109, 1.4219), z = c (1 + r) for the entry c of a 32-entry table (the top
// five fraction bits of z), |r| <= 2^-6, ln(1 + r) as its Taylor polynomial to
// r^9 (mc_ln_table.h, gen_ln_table.py); e^t by Cody-Waite reduction
// t = k ln2 + r, |r| <= ln2/2, and the Taylor polynomial of e^r to degree 12.
// Zero, negative, subnormal, infinite and NaN operands take the C99
// special-case results (subnormals a series logarithm) without a library call.
// Every multiply-add is an explicit fma and the table is a compile-time
// constant, so the host and the device produce the same bits. pow_f32 and
// log_f32 stay out of line: inlined at every call site of the adjoint they
// raised its register use from 128 to 169 VGPRs.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define DDRS_HD __host__ __device__
#define DDRS_TABLE __constant__ constexpr
#else
#define DDRS_HD
#define DDRS_TABLE constexpr
#endif

#include "mc_ln_table.h"

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

// ln(x) for finite x > 0 (double), ~2^-50 relative (absolute near x = 1), by the
// series 2 atanh(s), s = (m-1)/(m+1) for x = m 2^e, m in [1/sqrt2, sqrt2), to s^19:
// the logarithm of subnormal fp32 operands
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

DDRS_HD inline uint32_t f32_bits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}
DDRS_HD inline float bits_f32(uint32_t u) {
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// ln(x) for a positive normal finite fp32 x (double), ~2^-52 relative: the
// table reduction of the header comment. T is kLnTab or a copy of it (a kernel
// keeps one in LDS).
DDRS_HD inline __attribute__((always_inline)) double ln_tab(float x, const LnEnt* T) {
    const uint32_t u = f32_bits(x);
    const int32_t tmp = static_cast<int32_t>(u - kLnOff);
    const int i = (tmp >> (23 - kLnBits)) & (kLnEntries - 1);
    const int k = tmp >> 23;  // arithmetic shift: floor
    const float z = bits_f32(u - (static_cast<uint32_t>(tmp) & 0xff800000u));
    const LnEnt e = T[i];
    const double r = fma(static_cast<double>(z), e.invc, -1.0);  // |r| <= 2^-6
    double p = kc(1.0 / 9.0);
    p = fma(p, r, kc(-1.0 / 8.0));
    p = fma(p, r, kc(1.0 / 7.0));
    p = fma(p, r, kc(-1.0 / 6.0));
    p = fma(p, r, kc(1.0 / 5.0));
    p = fma(p, r, kc(-1.0 / 4.0));
    p = fma(p, r, kc(1.0 / 3.0));
    p = fma(p, r, -0.5);
    const double l1p = fma(p, r * r, r);  // ln(1 + r)
    const double ln2_hi = kc(6.93147180369123816490e-01);  // 21 trailing zero bits
    const double ln2_lo = kc(1.90821492927058770002e-10);
    const double dk = static_cast<double>(k);
    return fma(dk, ln2_hi, e.logc) + fma(dk, ln2_lo, l1p);
}

// whether x takes ln_tab: positive, normal and finite
DDRS_HD inline bool ln_tab_ok(float x) { return x >= 0x1p-126f && x < INFINITY; }

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

// y as an operand of this call: an exponent that is constant over a loop would
// otherwise have its special-case tests hoisted out of the loop as lane masks,
// which the register allocator then keeps in vector lanes and reads back with
// two lane reads per test and call (same value, same bits).
DDRS_HD inline float per_call(float y, float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(y) : "v"(x));
#else
    (void)x;
#endif
    return y;
}

// x^y for fp32 x, y, rounded once to fp32. T: kLnTab or a copy of it.
DDRS_HD inline __attribute__((always_inline)) float pow_f32_inl(float x, float y,
                                                                const LnEnt* T = kLnTab) {
    y = per_call(y, x);
    const double xd = static_cast<double>(x), yd = static_cast<double>(y);
    const bool fy = yd == yd && fabs(yd) < INFINITY;
    if (ln_tab_ok(x) && fy) return pow_from_ln(ln_tab(x, T), yd);
    if (pow_fast_path(xd, yd)) return pow_from_ln(ln_reduced(xd), yd);  // subnormal x
    return pow_special(xd, yd);
}
DDRS_HD inline DDRS_MATH_LINKAGE float pow_f32(float x, float y) { return pow_f32_inl(x, y); }

// ln x for fp32 x, rounded once to fp32
DDRS_HD inline __attribute__((always_inline)) float log_f32_inl(float x, const LnEnt* T = kLnTab) {
    if (ln_tab_ok(x)) return static_cast<float>(ln_tab(x, T));
    if (x > 0.0f && x < INFINITY)  // subnormal
        return static_cast<float>(ln_reduced(static_cast<double>(x)));
    return log_special(static_cast<double>(x));
}
DDRS_HD inline DDRS_MATH_LINKAGE float log_f32(float x) { return log_f32_inl(x); }

// pow_f32(x, y) and log_f32(x) from one evaluation of ln x; both results are
// bit-identical to the separate calls.
struct PowLn {
    float pw, ln;
};
DDRS_HD inline __attribute__((always_inline)) PowLn pow_ln_f32_inl(float x, float y,
                                                                   const LnEnt* T = kLnTab) {
    y = per_call(y, x);
    const double xd = static_cast<double>(x), yd = static_cast<double>(y);
    PowLn r;
    if (ln_tab_ok(x) && yd == yd && fabs(yd) < INFINITY) {
        const double l = ln_tab(x, T);
        r.ln = static_cast<float>(l);
        r.pw = pow_from_ln(l, yd);
    } else if (x > 0.0f && x < INFINITY) {  // subnormal x, or y NaN or infinite
        const double l = ln_tab_ok(x) ? ln_tab(x, T) : ln_reduced(xd);
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

// Per-reach Muskingum-Cunge physics for gfx950, held in registers.
//
// mc_coef:  S1..S23 of forward_chain_inner (src/routing/mmc_op.rs:1113-1236),
//           with the time-invariant S1/S3/S5/S15 terms precomputed per reach.
// mc_local_bwd: the per-reach part of timestep_backward_core B25..B1
//           (src/routing/mmc_op.rs:340-835); the two network terms (the
//           transposed-solve result gb and the N^T gi sum) come in as arguments.
//
// Every statement is one reference tensor op in the reference's evaluation
// order; the library is compiled with -ffp-contract=off so no multiply-add is
// fused and each op rounds to fp32 exactly once, as on the reference CPU path.
#pragma once

#include <hip/hip_runtime.h>

#include "mc_math.h"

namespace ddrs {

struct Phys {
    float dt, q_lb, depth_lb, v_lb, bw_lb;
};

// Variants of the physics (config.rs:580-672): default corrected physics,
// legacy ddr_match (celerity 5/3 v, X = x_storage), and enforce_positivity
// (S18' K floor, S19' X cap).
enum : int { PH_DEFAULT = 0, PH_DDR = 1, PH_POS = 2 };

constexpr float kDelta = 1e-2f;  // POSITIVITY_DELTA, mmc_op.rs:57

// x^y and ln(x) are evaluated in double and rounded once to fp32 (mc_math.h):
// on this path's inputs that reproduces the reference CPU's libm powf on 99.94%
// of calls, where the device library's fp32 powf agrees on only 79% (measured
// on gfx950); every disagreement seeds a rounding difference that the time
// loop amplifies.
__device__ __forceinline__ float rpow(float x, float y) { return pow_f32(x, y); }

__device__ __forceinline__ float fmaxr(float x, float lo) { return (x < lo) ? lo : x; }
__device__ __forceinline__ float fclamp(float x, float lo, float hi) {
    return (x < lo) ? lo : ((x > hi) ? hi : x);
}
__device__ __forceinline__ float fminr(float a, float b) { return (b < a) ? b : a; }

// Per-reach inputs that do not change over a window: the parameters and the
// reference's time-invariant subexpressions, evaluated once per window with
// the reference's own ops (so every step sees the bits it would compute).
struct Reach {
    float n, p, len, slope;         // slope after the slope_lb clamp
    float sqrt_s, xst, q_eps, qe1;  // sqrt(slope); x_storage; S1 q_spatial + 1e-6; S2 q_eps + 1
    float denominator, exponent;    // S3 p sqrt(S) + 1e-8; S5 (1 / (3 q_eps + 5)) * 3
    float inv_n, f53sq;             // S15 1 / n; B5 (3 q_eps + 5)^2
};

// Reach statics in HBM: three float4 per reach, schedule order.
__device__ __forceinline__ Reach load_reach(const float4* st, int p) {
    const float4 a = st[3 * p], b = st[3 * p + 1], c = st[3 * p + 2];
    Reach r;
    r.n = a.x; r.p = a.y; r.len = a.z; r.slope = a.w;
    r.sqrt_s = b.x; r.xst = b.y; r.q_eps = b.z; r.qe1 = b.w;
    r.denominator = c.x; r.exponent = c.y; r.inv_n = c.z; r.f53sq = c.w;
    return r;
}

__device__ __forceinline__ void make_reach(float n, float qsp, float psp, float len, float slope,
                                           float xst, float4* out) {
    const float sqrt_s = sqrtf(slope);
    const float q_eps = qsp + 1e-6f;
    const float f53 = q_eps * 3.0f + 5.0f;
    out[0] = make_float4(n, psp, len, slope);
    out[1] = make_float4(sqrt_s, xst, q_eps, q_eps + 1.0f);
    out[2] = make_float4(psp * sqrt_s + 1e-8f, (1.0f / f53) * 3.0f, 1.0f / n, f53 * f53);
}

// Correctly rounded quotients a / b of floats sharing one divisor b, from a
// reciprocal refined in double: y0 = rcp(b) (1 ulp), two Newton steps give
// r = 1/b within 2^-53 relative, and (double)a * r is within 2^-52 of a / b.
// The quotient of two floats lies at least 2^-49 (relative) from any fp32
// rounding midpoint, so rounding the double once gives the correctly rounded
// (IEEE) quotient: the same bits as a / b. That holds for a divisor of
// magnitude in [2^-100, 2^100] and a result that is zero or in the normal
// range (or overflows); div_ok() tells, per quotient, and the caller divides
// the IEEE way where it does not hold.
struct SharedDiv {
    double r;
    bool ok_b;
    __device__ __forceinline__ explicit SharedDiv(float b) {
        const float ab = fabsf(b);
        ok_b = ab >= 0x1p-100f && ab <= 0x1p100f;  // false for NaN
        const double bd = static_cast<double>(b);
        const double y0 = static_cast<double>(__builtin_amdgcn_rcpf(b));
        const double e0 = fma(-bd, y0, 1.0);
        const double y1 = fma(e0, y0, y0);
        const double e1 = fma(-bd, y1, 1.0);
        r = fma(e1, y1, y1);
    }
    __device__ __forceinline__ float div(float a) const {
        return static_cast<float>(static_cast<double>(a) * r);
    }
    __device__ __forceinline__ float inv() const { return static_cast<float>(r); }
    // q is zero, normal, infinite (overflow): not subnormal and not NaN
    static __device__ __forceinline__ bool div_ok(float q) {
        return fabsf(q) >= 0x1p-126f || q == 0.0f;
    }
};

struct Coef {
    float ratio, ln_ratio, depth, ln_depth, tw, ss_raw, ss, bw_raw, bw, hyd_r;
    float v_un, v_cl, cel, k_musk, x_eff, denom, c1, c2, c3, c4;
};

// The powers of one reach-step: S5 ratio^exponent and S7 depth^q_eps with the
// logarithms of their bases (B6, B7 use them), and S15 hyd_r^(2/3).
struct PowVals {
    float pw5, ln5, pw7, ln7, pw15;
};

// The power source of mc_coef: evaluates the three powers (out of line) and
// keeps their values.
struct PowEval {
    PowVals v;
    __device__ __forceinline__ PowLn p5(float x, float y) {
        const PowLn r = pow_ln_f32(x, y);
        v.pw5 = r.pw;
        v.ln5 = r.ln;
        return r;
    }
    __device__ __forceinline__ PowLn p7(float x, float y) {
        const PowLn r = pow_ln_f32(x, y);
        v.pw7 = r.pw;
        v.ln7 = r.ln;
        return r;
    }
    __device__ __forceinline__ float p15(float x) {
        v.pw15 = pow_f32(x, 2.0f / 3.0f);
        return v.pw15;
    }
};

// The same evaluations inlined: no call, so the caller's outstanding memory
// operations are not waited for at each power (a call waits for all of them).
// T: the logarithm table, kLnTab or a copy of it in LDS.
struct PowEvalInl {
    PowVals v;
    const LnEnt* T;
    __device__ explicit PowEvalInl(const LnEnt* tab) : T(tab) {}
    __device__ __forceinline__ PowLn p5(float x, float y) {
        const PowLn r = pow_ln_f32_inl(x, y, T);
        v.pw5 = r.pw;
        v.ln5 = r.ln;
        return r;
    }
    __device__ __forceinline__ PowLn p7(float x, float y) {
        const PowLn r = pow_ln_f32_inl(x, y, T);
        v.pw7 = r.pw;
        v.ln7 = r.ln;
        return r;
    }
    __device__ __forceinline__ float p15(float x) {
        v.pw15 = pow_f32_inl(x, 2.0f / 3.0f, T);
        return v.pw15;
    }
};

// S1..S23 for out[t] = q. ln_ratio and ln_depth are log_f32 of ratio and depth
// (needed by B6 and B7), taken from the same ln evaluation as the powers.
// SD: the four quotients over S20's denominator from one shared reciprocal
// (SharedDiv, bit-identical to the divisions); off, each is an IEEE division.
template <int PH, class Pow, bool SD = false>
__device__ __forceinline__ void mc_coef(const Phys& P, const Reach& r, float q, Coef& o,
                                        Pow& pw) {
    float numerator = (q * r.n) * r.qe1;                                // S2
    o.ratio = numerator / r.denominator;                                // S3, S4
    const PowLn d = pw.p5(o.ratio, r.exponent);                         // S5
    o.ln_ratio = d.ln;
    o.depth = fmaxr(d.pw, P.depth_lb);                                  // S6
    const PowLn w = pw.p7(o.depth, r.q_eps);
    o.ln_depth = w.ln;
    o.tw = r.p * w.pw;                                                  // S7
    o.ss_raw = (o.tw * r.q_eps) / (o.depth * 2.0f);                     // S8
    o.ss = fclamp(o.ss_raw, 0.5f, 50.0f);                               // S9
    o.bw_raw = o.tw - (o.ss * o.depth) * 2.0f;                          // S10
    o.bw = fmaxr(o.bw_raw, P.bw_lb);                                    // S11
    const float area = ((o.tw + o.bw) * o.depth) / 2.0f;                // S12
    const float root = sqrtf(o.ss * o.ss + 1.0f);
    const float wp = o.bw + (o.depth * root) * 2.0f;                    // S13
    o.hyd_r = area / wp;                                                // S14
    o.v_un = (r.inv_n * pw.p15(o.hyd_r)) * r.sqrt_s;                    // S15
    o.v_cl = fclamp(o.v_un, P.v_lb, 15.0f);                             // S16
    if (PH == PH_DDR) {                                                 // S17
        o.cel = o.v_cl * (5.0f / 3.0f);
    } else {
        float beta = ((-(area * root)) / (o.tw * wp)) * (4.0f / 3.0f) + (5.0f / 3.0f);
        o.cel = o.v_cl * beta;
    }
    float k_raw = r.len / o.cel;                                        // S18
    o.k_musk = (PH == PH_POS) ? fmaxr(k_raw, P.dt * (1.0f + kDelta) / 2.0f) : k_raw;
    if (PH == PH_DDR) {                                                 // S19
        o.x_eff = r.xst;
    } else {
        float wq = q / (((o.tw * r.slope) * o.cel) * r.len + 1e-12f);
        float x_cunge = fclamp(((-wq) + 1.0f) * 0.5f, 0.0f, 0.5f);
        if (PH == PH_POS) {
            float cr = (1.0f / o.k_musk) * P.dt;
            float hi_a = cr * (0.5f * (1.0f - kDelta));
            float hi_b = ((-cr) * 0.5f + 1.0f) * (1.0f - kDelta);
            o.x_eff = fminr(fminr(x_cunge, hi_a), hi_b);
        } else {
            o.x_eff = x_cunge;
        }
    }
    float one_minus_x = (-o.x_eff) + 1.0f;                              // S20..S23
    float two_k = o.k_musk * 2.0f;
    float two_kx = two_k * o.x_eff;
    float two_k_1mx = two_k * one_minus_x;
    o.denom = two_k_1mx + P.dt;
    const float n1 = (-two_kx) + P.dt;
    const float n2 = two_kx + P.dt;
    const float n3 = two_k_1mx - P.dt;
    float c1, c2, c3, inv;
    bool ieee = true;
    if (SD) {
        const SharedDiv sd(o.denom);
        c1 = sd.div(n1);
        c2 = sd.div(n2);
        c3 = sd.div(n3);
        inv = sd.inv();
        ieee = !(sd.ok_b && SharedDiv::div_ok(c1) && SharedDiv::div_ok(c2) &&
                 SharedDiv::div_ok(c3) && SharedDiv::div_ok(inv));
    }
    if (ieee) {
        c1 = n1 / o.denom;
        c2 = n2 / o.denom;
        c3 = n3 / o.denom;
        inv = 1.0f / o.denom;
    }
    o.c1 = c1;
    o.c2 = c2;
    o.c3 = c3;
    o.c4 = inv * (2.0f * P.dt);
}

struct LocalGrads {
    float gn, gqsp, gp, gq, gqp;
};

// B25..B1 for one reach. gb: this reach's transposed-solve value; gc1: the B26
// assemble-backward value; i_t: N q_t at this reach; gq24: (N^T gi) at this
// reach. Geometry is recomputed from the saved primitives exactly as the
// reference backward does (wp from S13's operands, area = R wp).
//
// Given the forward state, every output is linear in (gb, gc1, gq24): the
// clamp and branch masks depend on the forward state only, and gq24 enters gq
// alone. Since gc1 = gb * sum_k adj_k x[col k], the step's whole backward is
// gb times the call with gb = 1, gc1 = that sum, gq24 = 0, plus gq24 in gq.
// a / b: correctly rounded, or (F) as a times the hardware reciprocal (about
// 1 ulp), for values that only scale gradients.
template <bool F>
__device__ __forceinline__ float dv(float a, float b) {
    if (F) return a * __builtin_amdgcn_rcpf(b);
    return a / b;
}

// F = true evaluates the divisions that scale gradient values with the
// hardware reciprocal; the ones that decide clamp and branch masks stay
// correctly rounded, so the masks are those of the forward.
template <int PH, bool F = false>
__device__ __forceinline__ LocalGrads mc_local_bwd(const Phys& P, const Coef& c, const Reach& r,
                                                   float q, float qp, float i_t, float gb,
                                                   float gc1, float gq24) {
    LocalGrads g;
    const float dt = P.dt;
    const float nman = r.n, psp = r.p, length = r.len, slope = r.slope, sqrt_s = r.sqrt_s;
    // B25
    float gc2 = gb * i_t;
    float gc3 = gb * q;
    float gc4 = gb * qp;
    float gq25 = c.c3 * gb;
    g.gqp = c.c4 * gb;
    // B23..B20
    float denom_sq = c.denom * c.denom;
    float one_minus_x = (-c.x_eff) + 1.0f;
    float two_k = c.k_musk * 2.0f;
    float two_kx = two_k * c.x_eff;
    float two_k_1mx = two_k * one_minus_x;
    float num_c1 = (-two_kx) + dt;
    float num_c2 = two_kx + dt;
    float num_c3 = two_k_1mx - dt;
    float gd1 = dv<F>((-gc1) * num_c1, denom_sq);
    float gd2 = dv<F>((-gc2) * num_c2, denom_sq);
    float gd3 = dv<F>((-gc3) * num_c3, denom_sq);
    float gd4 = dv<F>((-gc4) * (2.0f * dt), denom_sq);
    float gdenom_total = ((gd1 + gd2) + gd3) + gd4;
    float g_2kx_total = dv<F>(-gc1, c.denom) + dv<F>(gc2, c.denom);
    float g_2k1mx_total = dv<F>(gc3, c.denom) + gdenom_total;
    // B19 / B19'
    float xg_q = 0.0f, xg_tw = 0.0f, xg_c = 0.0f, gk_cap = 0.0f;
    if (PH != PH_DDR) {
        float gxx = two_k * (g_2kx_total - g_2k1mx_total);
        float bscl = ((c.tw * slope) * c.cel) * length + 1e-12f;
        float w = q / bscl;
        float half_w = w * 0.5f;
        float x_for_clamp = c.x_eff;
        if (PH == PH_POS) {
            float x_cunge = fclamp(((-w) + 1.0f) * 0.5f, 0.0f, 0.5f);
            float cr = (1.0f / c.k_musk) * dt;
            float hi_a = cr * (0.5f * (1.0f - kDelta));
            float hi_b = ((-cr) * 0.5f + 1.0f) * (1.0f - kDelta);
            bool m_cunge = (x_cunge <= hi_a) && (x_cunge <= hi_b);
            bool m_a = (!m_cunge) && (hi_a <= hi_b);
            bool m_b = (!m_cunge) && (!m_a);
            float g_hi_a = m_a ? gxx : 0.0f;
            float g_hi_b = m_b ? gxx : 0.0f;
            float gx_cunge = m_cunge ? gxx : 0.0f;
            float gcr = (g_hi_a - g_hi_b) * (0.5f * (1.0f - kDelta));
            gk_cap = dv<F>((-gcr) * dt, c.k_musk * c.k_musk);
            gxx = gx_cunge;
            x_for_clamp = x_cunge;
        }
        bool unsat = (x_for_clamp > 0.0f) && (x_for_clamp < 0.5f);
        gxx = unsat ? gxx : 0.0f;
        xg_q = dv<F>((-gxx) * 0.5f, bscl);
        xg_tw = dv<F>(gxx * half_w, c.tw);
        xg_c = dv<F>(gxx * half_w, c.cel);
    }
    float g_2k_total = (g_2kx_total * c.x_eff) + (g_2k1mx_total * one_minus_x);
    float gk = g_2k_total * 2.0f;
    if (PH == PH_POS) gk = gk + gk_cap;
    // geometry recomputed from saved primitives
    float one_plus_ss_sq = c.ss * c.ss + 1.0f;
    float sq1 = sqrtf(one_plus_ss_sq);
    float wp = c.bw + (c.depth * sq1) * 2.0f;
    float area = c.hyd_r * wp;
    // B18'
    float gk_raw = gk;
    if (PH == PH_POS) {
        float k_raw = length / c.cel;
        float k_floor = dt * (1.0f + kDelta) / 2.0f;
        gk_raw = (k_raw > k_floor) ? gk : 0.0f;
    }
    // B18
    float gcel = dv<F>((-gk_raw) * length, c.cel * c.cel);
    if (PH != PH_DDR) gcel = gcel + xg_c;
    // B17
    float gvcl, bg_area = 0.0f, bg_tw = 0.0f, bg_wp = 0.0f, bg_ss = 0.0f;
    if (PH == PH_DDR) {
        gvcl = gcel * (5.0f / 3.0f);
    } else {
        float g_term = dv<F>(area * sq1, c.tw * wp) * (4.0f / 3.0f);
        float beta = (-g_term) + (5.0f / 3.0f);
        float gbeta = gcel * c.v_cl;
        bg_area = dv<F>((-gbeta) * g_term, area);
        bg_tw = dv<F>(gbeta * g_term, c.tw);
        bg_wp = dv<F>(gbeta * g_term, wp);
        bg_ss = dv<F>(((-gbeta) * g_term) * c.ss, one_plus_ss_sq);
        gvcl = gcel * beta;
    }
    // B16
    float gvun = ((c.v_un > P.v_lb) && (c.v_un < 15.0f)) ? gvcl : 0.0f;
    // B15
    float gn15 = gvun * dv<F>(-c.v_un, nman);
    float gr = dv<F>(gvun * (c.v_un * (2.0f / 3.0f)), c.hyd_r);
    // B14
    float garea_r = dv<F>(gr, wp);
    float gwp_r = dv<F>((-gr) * area, wp * wp);
    // B13
    float gwp = (PH == PH_DDR) ? gwp_r : (gwp_r + bg_wp);
    float gbw13 = gwp;
    float gd13 = (gwp * sq1) * 2.0f;
    float gss13 = dv<F>(((gwp * c.depth) * 2.0f) * c.ss, sq1);
    // B12
    float garea = (PH == PH_DDR) ? garea_r : (garea_r + bg_area);
    float half_d = c.depth * 0.5f;
    float gtw12 = garea * half_d;
    float gbw12 = garea * half_d;
    float gd12 = (garea * (c.tw + c.bw)) * 0.5f;
    // B11
    float gbw_total = gbw13 + gbw12;
    float gbw_raw = (c.bw_raw > P.bw_lb) ? gbw_total : 0.0f;
    // B10
    float gtw10 = gbw_raw;
    float gss10 = ((-gbw_raw) * 2.0f) * c.depth;
    float gd10 = ((-gbw_raw) * 2.0f) * c.ss;
    // B9
    float gss_comb = (PH == PH_DDR) ? (gss13 + gss10) : ((gss13 + gss10) + bg_ss);
    float gss_clamp = ((c.ss_raw > 0.5f) && (c.ss_raw < 50.0f)) ? gss_comb : 0.0f;
    // B8
    float two_d = c.depth * 2.0f;
    float gtw8 = dv<F>(gss_clamp * r.q_eps, two_d);
    float gqeps8 = dv<F>(gss_clamp * c.tw, two_d);
    float gd8 = dv<F>(((-gss_clamp) * c.tw) * r.q_eps, two_d * c.depth);
    float gtw_total = (gtw12 + gtw10) + gtw8;
    if (PH != PH_DDR) gtw_total = gtw_total + bg_tw;
    if (PH != PH_DDR) gtw_total = gtw_total + xg_tw;
    // B7
    float gp7 = dv<F>(gtw_total * c.tw, psp);
    float gdepth7 = dv<F>((gtw_total * c.tw) * r.q_eps, c.depth);
    float gqeps7 = (gtw_total * c.tw) * c.ln_depth;
    // B6
    float gd_total = (((gd13 + gd12) + gd10) + gd8) + gdepth7;
    float gd_pre = (c.depth > P.depth_lb) ? gd_total : 0.0f;
    float gratio6 = dv<F>((gd_pre * r.exponent) * c.depth, c.ratio);
    float gexp6 = (gd_pre * c.depth) * c.ln_ratio;
    // B5
    float gqeps5 = dv<F>((-gexp6) * 9.0f, r.f53sq);
    // B4
    float gnum = dv<F>(gratio6, r.denominator);
    float gden = dv<F>((-gratio6) * c.ratio, r.denominator);
    // B3
    float gp3 = gden * sqrt_s;
    // B2
    const float qe1 = r.qe1;
    float gq2 = (gnum * nman) * qe1;
    float gn2 = (gnum * q) * qe1;
    float gqeps2 = (gnum * q) * nman;
    // B1 and totals (mmc_op.rs:823-835)
    g.gqsp = ((gqeps8 + gqeps7) + gqeps5) + gqeps2;
    g.gn = gn15 + gn2;
    g.gp = gp7 + gp3;
    float gq = (gq25 + gq24) + gq2;
    if (PH != PH_DDR) gq = gq + xg_q;
    g.gq = gq;
    return g;
}

}  // namespace ddrs

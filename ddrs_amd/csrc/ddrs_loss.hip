// Gauge readout and training loss on gfx950:
//  - scatter_add_by_group (reference src/training/forward.rs:18-34): the
//    per-gauge sum of the routed discharge of the gauge's outflow reaches;
//  - tau_trim_and_downsample (src/training/loss.rs:33-95): hours
//    [tau, T - (24 - tau)) pooled into days with the area weights of
//    area_pool_weights, computed in float32 as the reference does;
//  - the batch losses (src/training/loss.rs:179-373: L1, nse-batch, nnse-kge,
//    kge) with the driver's warmup slice and NaN-gauge filter
//    (src/training/driver.rs:150-195);
//  - the three fused into the objective of one micro-batch, with its gradient
//    straight into dL/d(routed discharge).
// Everything is deterministic: sums run in a fixed order (the reference's k
// order for the gauge sums), reductions in double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

struct ddrs_gauge_plan {
    int N, G, K, U, max_t;
    int* gptr;     // [G+1] entries of gauge g: k in [gptr[g], gptr[g+1]) in increasing k
    int* gflat;    // [K] reach of each entry, gauge-major
    int* uptr;     // [U+1] entries of distinct reach u, in increasing k
    int* ureach;   // [U] reach id
    int* ugroup;   // [K] gauge of each entry, reach-major
    float* hourly;   // [G][max_t] scratch
    float* daily;    // [G][max_t / 24]
    float* gdaily;   // [G][max_t / 24]
    double* part;    // [G]
    int* valid;      // [G]
    int* n_valid;    // [1]
    double* gstats;  // [G][8] per-gauge moments of the KGE objectives
    float* gcoef;    // [G][6] their gradient coefficients and means
};

namespace ddrs {
namespace {

ddrs_status lfail(ddrs_status s, const std::string& m) {
    ddrs_internal_set_error(m.c_str());
    return s;
}

// The trimmed window and its day count (loss.rs:38-50).
bool window_of(int t_hours, int tau, int* start, int* L, int* M) {
    if (tau < 0 || tau >= 24) return false;
    *start = tau;
    const int end = t_hours - (24 - tau);
    if (end <= tau) return false;
    *L = end - tau;
    *M = *L / 24;
    return *M > 0;
}

// W(i, j) of area_pool_weights (loss.rs:70-95), in float32: the overlap of
// input cell j with output bin i over the bin width s; 0 outside the bin.
__device__ __forceinline__ float area_w(int i, int j, float s, int L) {
    const float left = static_cast<float>(i) * s;
    const float right = static_cast<float>(i + 1) * s;
    const int j_lo = static_cast<int>(floorf(left));
    const int j_hi = min(static_cast<int>(ceilf(right)), L);
    if (j < j_lo || j >= j_hi) return 0.0f;
    const float cl = fmaxf(static_cast<float>(j), left);
    const float cr = fminf(static_cast<float>(j + 1), right);
    return (cr - cl) / s;
}

// out[g][t] = sum_k runoff(flat_k, t) over gauge g's entries in k order.
__global__ void k_gauge_gather(const int* gptr, const int* gflat, const float* runoff, int T,
                               long long sn, long long st, float* out) {
    const int g = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    float s = 0.0f;
    for (int k = gptr[g]; k < gptr[g + 1]; ++k)
        s += runoff[static_cast<long long>(gflat[k]) * sn + static_cast<long long>(t) * st];
    out[static_cast<size_t>(g) * T + t] = s;
}

// grunoff(reach_u, t) = sum of gout[group][t] over reach u's entries in k order.
__global__ void k_gauge_scatter_bwd(const int* uptr, const int* ureach, const int* ugroup,
                                    const float* gout, int T, long long sn, long long st,
                                    float* grunoff) {
    const int u = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    float s = 0.0f;
    for (int k = uptr[u]; k < uptr[u + 1]; ++k) s += gout[static_cast<size_t>(ugroup[k]) * T + t];
    grunoff[static_cast<long long>(ureach[u]) * sn + static_cast<long long>(t) * st] = s;
}

// Row-wise forms for a time-major series (stride_t != 1): one workgroup per
// hour, threads across gauges (or reaches), so the reads of one hour share the
// row's pages instead of striding a whole row apart. The gauge series is kept
// hour-major, hourly[t][g].
__global__ void k_gauge_gather_rows(const int* gptr, const int* gflat, const float* runoff, int G,
                                    long long sn, long long st, float* hourly_tg) {
    const int t = blockIdx.x;
    const float* row = runoff + static_cast<long long>(t) * st;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        float s = 0.0f;
        for (int k = gptr[g]; k < gptr[g + 1]; ++k) s += row[static_cast<long long>(gflat[k]) * sn];
        hourly_tg[static_cast<size_t>(t) * G + g] = s;
    }
}

__global__ void k_gauge_scatter_bwd_rows(const int* uptr, const int* ureach, const int* ugroup,
                                         const float* g_tg, int G, int U, long long sn,
                                         long long st, float* grunoff) {
    const int t = blockIdx.x;
    const float* gr = g_tg + static_cast<size_t>(t) * G;
    float* row = grunoff + static_cast<long long>(t) * st;
    for (int u = threadIdx.x; u < U; u += blockDim.x) {
        float s = 0.0f;
        for (int k = uptr[u]; k < uptr[u + 1]; ++k) s += gr[ugroup[k]];
        row[static_cast<long long>(ureach[u]) * sn] = s;
    }
}

// [R][C] -> [C][R] through a 32 x 33 LDS tile.
__global__ void k_transpose(const float* in, int R, int C, float* out) {
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int dr = threadIdx.y; dr < 32; dr += blockDim.y) {
        const int r = r0 + dr, c = c0 + threadIdx.x;
        if (r < R && c < C) tile[dr][threadIdx.x] = in[static_cast<size_t>(r) * C + c];
    }
    __syncthreads();
    for (int dc = threadIdx.y; dc < 32; dc += blockDim.y) {
        const int c = c0 + dc, r = r0 + threadIdx.x;
        if (r < R && c < C) out[static_cast<size_t>(c) * R + r] = tile[threadIdx.x][dc];
    }
}

void transpose(const float* in, int R, int C, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_transpose, dim3((C + 31) / 32, (R + 31) / 32), dim3(32, 8), 0, s, in, R, C,
                       out);
}

// daily[g][i] from an hour-major series in_tg[t][g]: threads across gauges.
__global__ void k_area_pool_tg(const float* in_tg, int G, int start, int L, int M, float* out) {
    const int g = blockIdx.y * blockDim.x + threadIdx.x;
    const int i = blockIdx.x;
    if (g >= G) return;
    const float s = static_cast<float>(L) / static_cast<float>(M);
    const int j_lo = static_cast<int>(floorf(static_cast<float>(i) * s));
    const int j_hi = min(static_cast<int>(ceilf(static_cast<float>(i + 1) * s)), L);
    float acc = 0.0f;
    for (int j = j_lo; j < j_hi; ++j)
        acc = fmaf(in_tg[static_cast<size_t>(start + j) * G + g], area_w(i, j, s, L), acc);
    out[static_cast<size_t>(g) * M + i] = acc;
}

// gin_tg[h][g] = sum_i gout[g][i] W(i, h - start); zero outside the window.
__global__ void k_area_pool_bwd_tg(const float* gout, int G, int start, int L, int M,
                                   float* gin_tg) {
    const int g = blockIdx.y * blockDim.x + threadIdx.x;
    const int h = blockIdx.x;
    if (g >= G) return;
    const int j = h - start;
    float acc = 0.0f;
    if (j >= 0 && j < L) {
        const float s = static_cast<float>(L) / static_cast<float>(M);
        const int i0 = static_cast<int>(static_cast<float>(j) / s);
        for (int i = max(i0 - 1, 0); i <= min(i0 + 1, M - 1); ++i)
            acc = fmaf(gout[static_cast<size_t>(g) * M + i], area_w(i, j, s, L), acc);
    }
    gin_tg[static_cast<size_t>(h) * G + g] = acc;
}

// daily[g][i] = sum_j in[g][start + j] W(i, j), j ascending.
__global__ void k_area_pool(const float* in, int T, int start, int L, int M, float* out) {
    const int g = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float s = static_cast<float>(L) / static_cast<float>(M);
    const int j_lo = static_cast<int>(floorf(static_cast<float>(i) * s));
    const int j_hi = min(static_cast<int>(ceilf(static_cast<float>(i + 1) * s)), L);
    const float* x = in + static_cast<size_t>(g) * T + start;
    float acc = 0.0f;
    for (int j = j_lo; j < j_hi; ++j) acc = fmaf(x[j], area_w(i, j, s, L), acc);
    out[static_cast<size_t>(g) * M + i] = acc;
}

// gin[g][h] = sum_i gout[g][i] W(i, h - start); zero outside the window.
__global__ void k_area_pool_bwd(const float* gout, int T, int start, int L, int M, float* gin) {
    const int g = blockIdx.y;
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= T) return;
    const int j = h - start;
    float acc = 0.0f;
    if (j >= 0 && j < L) {
        const float s = static_cast<float>(L) / static_cast<float>(M);
        const int i0 = static_cast<int>(static_cast<float>(j) / s);
        for (int i = max(i0 - 1, 0); i <= min(i0 + 1, M - 1); ++i)
            acc = fmaf(gout[static_cast<size_t>(g) * M + i], area_w(i, j, s, L), acc);
    }
    gin[static_cast<size_t>(g) * T + h] = acc;
}

// Block sum of one double per thread, in a fixed tree order (256 threads).
__device__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Per gauge over days [t0, T): the sum of the loss terms, (p - o)^2 /
// (sigma + eps)^2 for nse-batch or |p - o| for L1, and whether the
// observations there hold no NaN (check_nan: else every gauge counts).
__global__ __launch_bounds__(256) void k_nse_partial(const float* p, const float* o,
                                                     const float* sigma, int T, int t0, float eps,
                                                     int check_nan, int kind, double* part,
                                                     int* valid) {
    __shared__ double sh[256];
    __shared__ int bad;
    const int g = blockIdx.x;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const bool l1 = kind == DDRS_LOSS_L1;
    const float d = l1 ? 1.0f : sigma[g] + eps;
    const float dd = d * d;
    double acc = 0.0;
    for (int t = t0 + threadIdx.x; t < T; t += blockDim.x) {
        const float ov = o[static_cast<size_t>(g) * T + t];
        if (check_nan && isnan(ov)) bad = 1;
        const float r = p[static_cast<size_t>(g) * T + t] - ov;
        acc += l1 ? static_cast<double>(fabsf(r)) : static_cast<double>(r * r / dd);
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        valid[g] = bad ? 0 : 1;
        part[g] = bad ? 0.0 : s;
    }
}

// loss = sum over kept gauges of part / (kept * days); n_valid = kept * days.
__global__ __launch_bounds__(256) void k_nse_final(const double* part, const int* valid, int G,
                                                   int days, float* loss, int* n_valid,
                                                   int* n_valid_out) {
    __shared__ double sh[256];
    double acc = 0.0, cnt = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        acc += part[g];
        cnt += valid[g];
    }
    const double s = block_sum(acc, sh);
    const double c = block_sum(cnt, sh);
    if (threadIdx.x == 0) {
        const double n = c * days;
        if (loss) *loss = n > 0 ? static_cast<float>(s / n) : 0.0f;
        if (n_valid) *n_valid = static_cast<int>(n);
        if (n_valid_out) *n_valid_out = static_cast<int>(n);
    }
}

// gp[g][t] = gloss * 2 (p - o) / ((sigma + eps)^2 n) (nse-batch) or
// gloss * sign(p - o) / n (L1, sign 0 at p == o) for kept gauges and t >= t0,
// else 0; n from *n_valid (or G * (T - t0) when n_valid is null).
__global__ void k_nse_bwd(const float* p, const float* o, const float* sigma, int G, int T,
                          int t0, float eps, int kind, const float* gloss, const int* valid,
                          const int* n_valid, float* gp) {
    const int g = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const size_t e = static_cast<size_t>(g) * T + t;
    float v = 0.0f;
    const int n = n_valid ? *n_valid : G * (T - t0);
    if (t >= t0 && n > 0 && (!valid || valid[g])) {
        const float gl = gloss ? *gloss : 1.0f;
        const float r = p[e] - o[e];
        if (kind == DDRS_LOSS_L1) {
            const float sg = (r > 0.0f) ? 1.0f : ((r < 0.0f) ? -1.0f : 0.0f);
            v = gl * sg / static_cast<float>(n);
        } else {
            const float d = sigma[g] + eps;
            v = gl * (2.0f * r / (d * d)) / static_cast<float>(n);
        }
    }
    gp[e] = v;
}

// The per-gauge objectives nnse-kge and kge (loss.rs:257-373) over days
// [t0, T). Moments of gauge g, in double: st[g] = {mean p, mean o, population
// variances of p and o, covariance, sum of squared residuals, sum of squared
// centred o}; valid[g] = no NaN in its observations there.
__global__ __launch_bounds__(256) void k_kge_stats(const float* p, const float* o, int T, int t0,
                                                   double* st, int* valid) {
    __shared__ double sh[256];
    __shared__ int bad;
    const int g = blockIdx.x;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const float* pg = p + static_cast<size_t>(g) * T;
    const float* og = o + static_cast<size_t>(g) * T;
    double sp = 0.0, so = 0.0;
    for (int t = t0 + threadIdx.x; t < T; t += blockDim.x) {
        const float ov = og[t];
        if (isnan(ov)) bad = 1;
        sp += pg[t];
        so += ov;
    }
    const double n = static_cast<double>(T - t0);
    const double mp = block_sum(sp, sh) / n;
    const double mo = block_sum(so, sh) / n;
    double vp = 0.0, vo = 0.0, cv = 0.0, se = 0.0;
    for (int t = t0 + threadIdx.x; t < T; t += blockDim.x) {
        const double pc = pg[t] - mp, oc = og[t] - mo;
        const double r = static_cast<double>(pg[t]) - static_cast<double>(og[t]);
        vp += pc * pc;
        vo += oc * oc;
        cv += pc * oc;
        se += r * r;
    }
    vp = block_sum(vp, sh);
    vo = block_sum(vo, sh);
    cv = block_sum(cv, sh);
    se = block_sum(se, sh);
    if (threadIdx.x == 0) {
        double* s = st + static_cast<size_t>(g) * 8;
        s[0] = mp;
        s[1] = mo;
        s[2] = vp / n;
        s[3] = vo / n;
        s[4] = cv / n;
        s[5] = se;
        s[6] = vo;
        valid[g] = bad ? 0 : 1;
    }
}

// Loss (mean over kept gauges) and, per kept gauge, the coefficients of
// dL/dp_t = A (p_t - mean p) + B (o_t - mean o) + C (p_t - o_t) + D, from
// d r = (o_t - mean o) / (n sp so) - r (p_t - mean p) / (n sp^2),
// d alpha = (p_t - mean p) / (n sp so), d beta = 1 / (n (mean o + eps)) and
// d (1 - NNSE) = 2 NNSE^2 (p_t - o_t) / sso (sp, so: the std's with eps).
__global__ __launch_bounds__(256) void k_kge_final(const double* st, const int* valid, int G,
                                                   int days, ddrs_loss_config cfg, float* loss,
                                                   int* n_valid, int* n_valid_out,
                                                   float* coef) {
    __shared__ double sh[256];
    double cnt = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) cnt += valid[g];
    const double kept = block_sum(cnt, sh);
    const double n = static_cast<double>(days), eps = cfg.eps;
    double acc = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        float* c = coef + static_cast<size_t>(g) * 6;
        const double* s = st + static_cast<size_t>(g) * 8;
        if (!valid[g]) {
            for (int i = 0; i < 6; ++i) c[i] = 0.0f;
            continue;
        }
        const double mp = s[0], mo = s[1];
        const double sp = sqrt(s[2] + eps), so = sqrt(s[3] + eps);
        const double r = s[4] / (sp * so), alpha = sp / so, beta = mp / (mo + eps);
        const double sso = s[6] + eps;
        const double nnse = 1.0 / (1.0 + s[5] / sso);
        const double dr = r - 1.0, da = alpha - 1.0, db = beta - 1.0;
        const double jr_pc = -r / (n * sp * sp), jr_oc = 1.0 / (n * sp * so);
        const double ja_pc = 1.0 / (n * sp * so), jb = 1.0 / (n * (mo + eps));
        const double g_resid = 2.0 * nnse * nnse / sso;
        double lg, A, B, C, D;
        if (cfg.kind == DDRS_LOSS_NNSE_KGE) {
            const double dist = sqrt(dr * dr + da * da + db * db);
            lg = cfg.nnse_weight * (1.0 - nnse) + cfg.kge_weight * dist;
            const double w = cfg.kge_weight / dist;
            A = w * (dr * jr_pc + da * ja_pc);
            B = w * dr * jr_oc;
            D = w * db * jb;
            C = cfg.nnse_weight * g_resid;
        } else {
            const double term = cfg.r_weight * dr * dr + cfg.alpha_weight * da * da +
                                cfg.beta_weight * db * db;
            const double m = (term <= cfg.kge_clamp) ? 1.0 : 0.0;
            lg = fmin(term, static_cast<double>(cfg.kge_clamp));
            A = 2.0 * m * (cfg.r_weight * dr * jr_pc + cfg.alpha_weight * da * ja_pc);
            B = 2.0 * m * cfg.r_weight * dr * jr_oc;
            D = 2.0 * m * cfg.beta_weight * db * jb;
            C = 0.0;
            if (cfg.nnse_weight != 0.0f) {
                lg += cfg.nnse_weight * (1.0 - nnse);
                C = cfg.nnse_weight * g_resid;
            }
        }
        acc += lg;
        c[0] = static_cast<float>(A / kept);
        c[1] = static_cast<float>(B / kept);
        c[2] = static_cast<float>(C / kept);
        c[3] = static_cast<float>(D / kept);
        c[4] = static_cast<float>(mp);
        c[5] = static_cast<float>(mo);
    }
    const double total = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        if (loss) *loss = kept > 0 ? static_cast<float>(total / kept) : 0.0f;
        if (n_valid) *n_valid = static_cast<int>(kept);
        if (n_valid_out) *n_valid_out = static_cast<int>(kept);
    }
}

__global__ void k_kge_bwd(const float* p, const float* o, int G, int T, int t0, const int* valid,
                          const float* coef, float* gp) {
    const int g = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const size_t e = static_cast<size_t>(g) * T + t;
    float v = 0.0f;
    if (t >= t0 && valid[g]) {
        const float* c = coef + static_cast<size_t>(g) * 6;
        v = c[0] * (p[e] - c[4]) + c[1] * (o[e] - c[5]) + c[2] * (p[e] - o[e]) + c[3];
    }
    gp[e] = v;
}

dim3 grid2(int n, int rows) { return dim3((n + 255) / 256, rows); }

ddrs_status hip_check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lfail(DDRS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return DDRS_OK;
}

}  // namespace
}  // namespace ddrs

using namespace ddrs;

extern "C" {

ddrs_status ddrs_gauge_plan_create(int32_t n_reaches, int32_t num_gauges, int32_t K,
                                   const int32_t* h_flat_indices, const int32_t* h_group_ids,
                                   int32_t max_t_hours, ddrs_gauge_plan_t* out) {
    if (!out) return lfail(DDRS_ERR_INVALID, "gauge plan: NULL out");
    *out = nullptr;
    if (n_reaches < 1 || num_gauges < 1 || K < 0 || max_t_hours < 1 ||
        (K > 0 && (!h_flat_indices || !h_group_ids)))
        return lfail(DDRS_ERR_INVALID, "gauge plan: bad sizes or NULL index arrays");
    if (num_gauges > 65535)
        return lfail(DDRS_ERR_CAPACITY, "gauge plan: at most 65535 gauges in this build");
    for (int k = 0; k < K; ++k) {
        if (h_flat_indices[k] < 0 || h_flat_indices[k] >= n_reaches)
            return lfail(DDRS_ERR_INVALID, "gauge plan: flat index " + std::to_string(k) +
                                               " outside [0, n_reaches)");
        if (h_group_ids[k] < 0 || h_group_ids[k] >= num_gauges)
            return lfail(DDRS_ERR_INVALID, "gauge plan: group id " + std::to_string(k) +
                                               " outside [0, num_gauges)");
    }
    // stable CSR by gauge and by reach (entries keep increasing k)
    std::vector<int> gptr(num_gauges + 1, 0), gflat(K);
    for (int k = 0; k < K; ++k) ++gptr[h_group_ids[k] + 1];
    for (int g = 0; g < num_gauges; ++g) gptr[g + 1] += gptr[g];
    {
        std::vector<int> fill(gptr.begin(), gptr.end() - 1);
        for (int k = 0; k < K; ++k) gflat[fill[h_group_ids[k]]++] = h_flat_indices[k];
    }
    std::vector<int> order(K);
    for (int k = 0; k < K; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return h_flat_indices[a] < h_flat_indices[b]; });
    std::vector<int> uptr(1, 0), ureach, ugroup(K);
    for (int q = 0; q < K; ++q) {
        const int k = order[q];
        if (q == 0 || h_flat_indices[k] != h_flat_indices[order[q - 1]]) {
            ureach.push_back(h_flat_indices[k]);
            if (q > 0) uptr.push_back(q);
        }
        ugroup[q] = h_group_ids[k];
    }
    if (K > 0) uptr.push_back(K);
    if (ureach.size() > 65535)
        return lfail(DDRS_ERR_CAPACITY, "gauge plan: at most 65535 distinct outflow reaches");
    auto* p = new (std::nothrow) ddrs_gauge_plan{};
    if (!p) return lfail(DDRS_ERR_NOMEM, "gauge plan: host allocation");
    p->N = n_reaches;
    p->G = num_gauges;
    p->K = K;
    p->U = static_cast<int>(ureach.size());
    p->max_t = max_t_hours;
    const size_t G = num_gauges, T = max_t_hours, D = std::max(1, max_t_hours / 24);
    bool ok = true;
    auto dalloc = [&](void** ptr, size_t bytes) {
        if (ok && hipMalloc(ptr, std::max<size_t>(bytes, 4)) != hipSuccess) ok = false;
    };
    dalloc(reinterpret_cast<void**>(&p->gptr), (G + 1) * 4);
    dalloc(reinterpret_cast<void**>(&p->gflat), static_cast<size_t>(K) * 4);
    dalloc(reinterpret_cast<void**>(&p->uptr), (ureach.size() + 1) * 4);
    dalloc(reinterpret_cast<void**>(&p->ureach), ureach.size() * 4);
    dalloc(reinterpret_cast<void**>(&p->ugroup), static_cast<size_t>(K) * 4);
    dalloc(reinterpret_cast<void**>(&p->hourly), G * T * 4);
    dalloc(reinterpret_cast<void**>(&p->daily), G * D * 4);
    dalloc(reinterpret_cast<void**>(&p->gdaily), G * D * 4);
    dalloc(reinterpret_cast<void**>(&p->part), G * 8);
    dalloc(reinterpret_cast<void**>(&p->valid), G * 4);
    dalloc(reinterpret_cast<void**>(&p->n_valid), 4);
    dalloc(reinterpret_cast<void**>(&p->gstats), G * 8 * 8);
    dalloc(reinterpret_cast<void**>(&p->gcoef), G * 6 * 4);
    auto up = [&](void* d, const void* h, size_t bytes) {
        if (ok && bytes && hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) ok = false;
    };
    up(p->gptr, gptr.data(), gptr.size() * 4);
    up(p->gflat, gflat.data(), gflat.size() * 4);
    up(p->uptr, uptr.data(), uptr.size() * 4);
    up(p->ureach, ureach.data(), ureach.size() * 4);
    up(p->ugroup, ugroup.data(), ugroup.size() * 4);
    if (!ok) {
        ddrs_gauge_plan_destroy(p);
        return lfail(DDRS_ERR_NOMEM, "gauge plan: device allocation or copy failed");
    }
    *out = p;
    return DDRS_OK;
}

void ddrs_gauge_plan_destroy(ddrs_gauge_plan_t p) {
    if (!p) return;
    for (void* q : {static_cast<void*>(p->gptr), static_cast<void*>(p->gflat),
                    static_cast<void*>(p->uptr), static_cast<void*>(p->ureach),
                    static_cast<void*>(p->ugroup), static_cast<void*>(p->hourly),
                    static_cast<void*>(p->daily), static_cast<void*>(p->gdaily),
                    static_cast<void*>(p->part), static_cast<void*>(p->valid),
                    static_cast<void*>(p->gstats), static_cast<void*>(p->gcoef),
                    static_cast<void*>(p->n_valid)})
        if (q) (void)hipFree(q);
    delete p;
}

ddrs_status ddrs_scatter_add_by_group(ddrs_gauge_plan_t p, const float* d_runoff, int32_t T,
                                      int64_t stride_n, int64_t stride_t, float* d_out,
                                      void* stream) {
    if (!p || T < 0 || (T > 0 && (!d_runoff || !d_out)))
        return lfail(DDRS_ERR_INVALID, "scatter_add_by_group: bad arguments");
    if (T == 0) return DDRS_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (stride_t == 1) {
        hipLaunchKernelGGL(k_gauge_gather, grid2(T, p->G), dim3(256), 0, s, p->gptr, p->gflat,
                           d_runoff, T, static_cast<long long>(stride_n), 1LL, d_out);
    } else {
        if (T > p->max_t)
            return lfail(DDRS_ERR_CAPACITY, "scatter_add_by_group: T exceeds the plan's max_t_hours");
        hipLaunchKernelGGL(k_gauge_gather_rows, dim3(T), dim3(256), 0, s, p->gptr, p->gflat,
                           d_runoff, p->G, static_cast<long long>(stride_n),
                           static_cast<long long>(stride_t), p->hourly);
        transpose(p->hourly, T, p->G, d_out, s);
    }
    return hip_check("scatter_add_by_group");
}

ddrs_status ddrs_scatter_add_by_group_bwd(ddrs_gauge_plan_t p, const float* d_gout, int32_t T,
                                          int64_t stride_n, int64_t stride_t, float* d_grunoff,
                                          void* stream) {
    if (!p || T < 0 || (T > 0 && (!d_gout || !d_grunoff)))
        return lfail(DDRS_ERR_INVALID, "scatter_add_by_group_bwd: bad arguments");
    if (T == 0 || p->U == 0) return DDRS_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (stride_t == 1) {
        hipLaunchKernelGGL(k_gauge_scatter_bwd, grid2(T, p->U), dim3(256), 0, s, p->uptr,
                           p->ureach, p->ugroup, d_gout, T, static_cast<long long>(stride_n), 1LL,
                           d_grunoff);
    } else {
        if (T > p->max_t)
            return lfail(DDRS_ERR_CAPACITY,
                         "scatter_add_by_group_bwd: T exceeds the plan's max_t_hours");
        transpose(d_gout, p->G, T, p->hourly, s);
        hipLaunchKernelGGL(k_gauge_scatter_bwd_rows, dim3(T), dim3(256), 0, s, p->uptr, p->ureach,
                           p->ugroup, p->hourly, p->G, p->U, static_cast<long long>(stride_n),
                           static_cast<long long>(stride_t), d_grunoff);
    }
    return hip_check("scatter_add_by_group_bwd");
}

ddrs_status ddrs_tau_trim_downsample(const float* d_in, int32_t G, int32_t T_hours, int32_t tau,
                                     float* d_out, void* stream) {
    int start, L, M;
    if (G < 0 || !window_of(T_hours, tau, &start, &L, &M))
        return lfail(DDRS_ERR_INVALID, "tau_trim_downsample: tau must be in [0, 24) and "
                                       "T_hours - 24 >= 24");
    if (G == 0) return DDRS_OK;
    if (!d_in || !d_out) return lfail(DDRS_ERR_INVALID, "tau_trim_downsample: NULL buffer");
    hipLaunchKernelGGL(k_area_pool, grid2(M, G), dim3(256), 0, static_cast<hipStream_t>(stream),
                       d_in, T_hours, start, L, M, d_out);
    return hip_check("tau_trim_downsample");
}

ddrs_status ddrs_tau_trim_downsample_bwd(const float* d_gout, int32_t G, int32_t T_hours,
                                         int32_t tau, float* d_gin, void* stream) {
    int start, L, M;
    if (G < 0 || !window_of(T_hours, tau, &start, &L, &M))
        return lfail(DDRS_ERR_INVALID, "tau_trim_downsample_bwd: tau must be in [0, 24) and "
                                       "T_hours - 24 >= 24");
    if (G == 0) return DDRS_OK;
    if (!d_gout || !d_gin) return lfail(DDRS_ERR_INVALID, "tau_trim_downsample_bwd: NULL buffer");
    hipLaunchKernelGGL(k_area_pool_bwd, grid2(T_hours, G), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d_gout, T_hours, start, L, M, d_gin);
    return hip_check("tau_trim_downsample_bwd");
}

ddrs_status ddrs_nse_batch_loss(ddrs_gauge_plan_t p, const float* d_p, const float* d_o,
                                const float* d_sigma, int32_t G, int32_t T, float eps,
                                float* d_loss, void* stream) {
    if (!p || G < 1 || G > p->G || T < 1 || !d_p || !d_o || !d_sigma || !d_loss)
        return lfail(DDRS_ERR_INVALID, "nse_batch_loss: bad arguments (G <= the plan's gauges)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_nse_partial, dim3(G), dim3(256), 0, s, d_p, d_o, d_sigma, T, 0, eps, 0,
                       static_cast<int>(DDRS_LOSS_NSE_BATCH), p->part, p->valid);
    hipLaunchKernelGGL(k_nse_final, dim3(1), dim3(256), 0, s, p->part, p->valid, G, T, d_loss,
                       nullptr, nullptr);
    return hip_check("nse_batch_loss");
}

ddrs_status ddrs_nse_batch_loss_bwd(const float* d_p, const float* d_o, const float* d_sigma,
                                    int32_t G, int32_t T, float eps, const float* d_gloss,
                                    float* d_gp, void* stream) {
    if (G < 1 || T < 1 || !d_p || !d_o || !d_sigma || !d_gp)
        return lfail(DDRS_ERR_INVALID, "nse_batch_loss_bwd: bad arguments");
    hipLaunchKernelGGL(k_nse_bwd, grid2(T, G), dim3(256), 0, static_cast<hipStream_t>(stream),
                       d_p, d_o, d_sigma, G, T, 0, eps, static_cast<int>(DDRS_LOSS_NSE_BATCH),
                       d_gloss, nullptr, nullptr, d_gp);
    return hip_check("nse_batch_loss_bwd");
}

ddrs_status ddrs_gauge_nse_loss(ddrs_gauge_plan_t p, const float* d_runoff, int32_t T_hours,
                                int64_t stride_n, int64_t stride_t, int32_t tau, int32_t warmup,
                                const float* d_obs, const float* d_sigma, float eps,
                                float* d_loss, int32_t* d_n_valid, float* d_grunoff,
                                float* d_hourly, void* stream) {
    if (!d_sigma) return lfail(DDRS_ERR_INVALID, "gauge_nse_loss: NULL sigma");
    ddrs_loss_config cfg{};
    cfg.kind = DDRS_LOSS_NSE_BATCH;
    cfg.eps = eps;
    return ddrs_gauge_batch_loss(p, &cfg, d_runoff, T_hours, stride_n, stride_t, tau, warmup,
                                 d_obs, d_sigma, d_loss, d_n_valid, d_grunoff, d_hourly, stream);
}

ddrs_status ddrs_gauge_batch_loss(ddrs_gauge_plan_t p, const ddrs_loss_config* cfg,
                                  const float* d_runoff, int32_t T_hours, int64_t stride_n,
                                  int64_t stride_t, int32_t tau, int32_t warmup,
                                  const float* d_obs, const float* d_sigma, float* d_loss,
                                  int32_t* d_n_valid, float* d_grunoff, float* d_hourly,
                                  void* stream) {
    int start, L, M;
    if (!cfg) return lfail(DDRS_ERR_INVALID, "gauge_batch_loss: NULL loss config");
    const int kind = cfg->kind;
    const float eps = cfg->eps;
    const bool per_gauge = kind == DDRS_LOSS_NNSE_KGE || kind == DDRS_LOSS_KGE;
    if (kind != DDRS_LOSS_L1 && kind != DDRS_LOSS_NSE_BATCH && !per_gauge)
        return lfail(DDRS_ERR_INVALID, "gauge_batch_loss: unknown loss kind");
    if (!p || !d_runoff || !d_obs || !d_loss || (kind == DDRS_LOSS_NSE_BATCH && !d_sigma))
        return lfail(DDRS_ERR_INVALID, "gauge_batch_loss: NULL argument");
    if (T_hours > p->max_t)
        return lfail(DDRS_ERR_CAPACITY, "gauge_batch_loss: T_hours exceeds the plan's max_t_hours");
    if (!window_of(T_hours, tau, &start, &L, &M))
        return lfail(DDRS_ERR_INVALID, "gauge_batch_loss: tau must be in [0, 24) and "
                                       "T_hours - 24 >= 24");
    if (warmup < 0 || warmup >= M)
        return lfail(DDRS_ERR_INVALID, "gauge_batch_loss: warmup must be in [0, days)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int G = p->G;
    const bool rows = stride_t != 1;  // time-major: hour-major gauge series
    if (rows) {
        hipLaunchKernelGGL(k_gauge_gather_rows, dim3(T_hours), dim3(256), 0, s, p->gptr, p->gflat,
                           d_runoff, G, static_cast<long long>(stride_n),
                           static_cast<long long>(stride_t), p->hourly);
        if (d_hourly) transpose(p->hourly, T_hours, G, d_hourly, s);
        hipLaunchKernelGGL(k_area_pool_tg, dim3(M, (G + 255) / 256), dim3(256), 0, s, p->hourly,
                           G, start, L, M, p->daily);
    } else {
        float* hourly = d_hourly ? d_hourly : p->hourly;
        hipLaunchKernelGGL(k_gauge_gather, grid2(T_hours, G), dim3(256), 0, s, p->gptr, p->gflat,
                           d_runoff, T_hours, static_cast<long long>(stride_n), 1LL, hourly);
        hipLaunchKernelGGL(k_area_pool, grid2(M, G), dim3(256), 0, s, hourly, T_hours, start, L,
                           M, p->daily);
    }
    if (per_gauge) {
        hipLaunchKernelGGL(k_kge_stats, dim3(G), dim3(256), 0, s, p->daily, d_obs, M, warmup,
                           p->gstats, p->valid);
        hipLaunchKernelGGL(k_kge_final, dim3(1), dim3(256), 0, s, p->gstats, p->valid, G,
                           M - warmup, *cfg, d_loss, p->n_valid, d_n_valid, p->gcoef);
    } else {
        hipLaunchKernelGGL(k_nse_partial, dim3(G), dim3(256), 0, s, p->daily, d_obs, d_sigma, M,
                           warmup, eps, 1, kind, p->part, p->valid);
        hipLaunchKernelGGL(k_nse_final, dim3(1), dim3(256), 0, s, p->part, p->valid, G,
                           M - warmup, d_loss, p->n_valid, d_n_valid);
    }
    if (d_grunoff && p->U > 0) {
        if (per_gauge)
            hipLaunchKernelGGL(k_kge_bwd, grid2(M, G), dim3(256), 0, s, p->daily, d_obs, G, M,
                               warmup, p->valid, p->gcoef, p->gdaily);
        else
            hipLaunchKernelGGL(k_nse_bwd, grid2(M, G), dim3(256), 0, s, p->daily, d_obs, d_sigma,
                               G, M, warmup, eps, kind, nullptr, p->valid, p->n_valid,
                               p->gdaily);
        if (rows) {
            hipLaunchKernelGGL(k_area_pool_bwd_tg, dim3(T_hours, (G + 255) / 256), dim3(256), 0, s,
                               p->gdaily, G, start, L, M, p->hourly);
            hipLaunchKernelGGL(k_gauge_scatter_bwd_rows, dim3(T_hours), dim3(256), 0, s, p->uptr,
                               p->ureach, p->ugroup, p->hourly, G, p->U,
                               static_cast<long long>(stride_n), static_cast<long long>(stride_t),
                               d_grunoff);
        } else {
            hipLaunchKernelGGL(k_area_pool_bwd, grid2(T_hours, G), dim3(256), 0, s, p->gdaily,
                               T_hours, start, L, M, p->hourly);
            hipLaunchKernelGGL(k_gauge_scatter_bwd, grid2(T_hours, p->U), dim3(256), 0, s,
                               p->uptr, p->ureach, p->ugroup, p->hourly, T_hours,
                               static_cast<long long>(stride_n), 1LL, d_grunoff);
        }
    }
    return hip_check("gauge_batch_loss");
}

}  // extern "C"

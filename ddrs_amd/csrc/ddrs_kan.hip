// KAN parameter head on gfx950: forward and reverse mode of
//   h0 = inputs W_in^T + b_in,  h_{l+1} = KanLayer_l(h_l),  out = sigmoid(h_L W_out^T + b_out)
// (reference src/nn/kan_head.rs:187-215). KanLayer follows pykan's KANLayer,
// which the reference's rskan v0.1.3 restates:
//   y[o] = sum_i mask[i,o] (scale_base[i,o] silu(x_i) + scale_sp[i,o] sum_c coef[i,o,c] B_c(x_i))
// with B_c the order-k B-spline basis on the knots of input i (Cox-de Boor).
//
// One thread per reach. A reach's input falls in one knot span, so only the k+1
// basis functions of that span are evaluated, by the same recursion restricted
// to them (the others are exactly zero). Per-thread vectors (activations,
// gradients) live in LDS, one column per lane. Parameter gradients are summed
// over the block's reaches in LDS and added to global memory once per block;
// the coefficient gradients go straight to global memory when the block's LDS
// cannot hold them. Float atomics make the gradient sums order-dependent at
// the last bits.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

namespace ddrs {
namespace {

constexpr int kKanMaxK = 5;       // spline order
constexpr int kKanMaxH = 64;      // hidden width
constexpr int kKanMaxP = 16;      // outputs
constexpr int kKanMaxL = 8;       // KAN layers
constexpr int kKanLanes = 64;     // threads per workgroup (one wave)
constexpr size_t kKanLds = 65536; // LDS budget of a workgroup

struct KanDev {
    int F, H, P, L, G, k, C;
    ddrs_kan_params p;
};

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

// Knot span of x on knots g[0..G): j with g[j] <= x < g[j+1], or -1 outside
// [g[0], g[G-1]) (all basis functions are zero there).
__device__ __forceinline__ int knot_span(const float* g, int G, float x) {
    if (!(x >= g[0] && x < g[G - 1])) return -1;
    int lo = 0, hi = G - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The k+1 basis functions that can be non-zero at x, B[t] = B_{j-k+t}, and
// their derivatives; Cox-de Boor with the same operations as the full
// recursion. At order kk the basis has G-1-kk functions (c <= G-2-kk); the
// knots are strictly ascending.
__device__ __forceinline__ void span_basis(const float* g, int G, int k, int j, float x,
                                           float (&B)[kKanMaxK + 1], float (&dB)[kKanMaxK + 1]) {
#pragma unroll
    for (int t = 0; t <= kKanMaxK; ++t) {
        B[t] = 0.0f;
        dB[t] = 0.0f;
    }
    B[k] = 1.0f;
    for (int kk = 1; kk <= k; ++kk) {
#pragma unroll
        for (int t = 0; t <= kKanMaxK; ++t) {
            if (t <= k) {
                const int c = j - k + t;
                const float b0 = B[t];
                const float b1 = (t < k) ? B[t + 1] : 0.0f;
                const float db0 = dB[t];
                const float db1 = (t < k) ? dB[t + 1] : 0.0f;
                float v = 0.0f, dv = 0.0f;
                if (c >= 0 && c <= G - 2 - kk && t >= k - kk) {
                    const float d1 = g[c + kk] - g[c];
                    const float d2 = g[c + kk + 1] - g[c + 1];
                    const float a = (x - g[c]) / d1;
                    const float cc = (g[c + kk + 1] - x) / d2;
                    v = a * b0 + cc * b1;
                    dv = ((b0 / d1 + a * db0) - b1 / d2) + cc * db1;
                }
                B[t] = v;
                dB[t] = dv;
            }
        }
    }
}

// The head's forward for the reach of this lane; writes the input of layer
// `stop` (stop == L: the last activation) into column `lane` of xs[H][64].
__device__ void kan_forward_to(const KanDev& d, const float* in, int stop, float* xs, float* ys,
                               int lane) {
    const int H = d.H, F = d.F;
    for (int j = 0; j < H; ++j) {
        float acc = 0.0f;
        for (int f = 0; f < F; ++f) acc = acc + in[f] * d.p.input_weight[j * F + f];
        xs[j * kKanLanes + lane] = acc + d.p.input_bias[j];
    }
    for (int l = 0; l < stop; ++l) {
        const size_t ll = static_cast<size_t>(l) * H * H;
        for (int o = 0; o < H; ++o) ys[o * kKanLanes + lane] = 0.0f;
        for (int i = 0; i < H; ++i) {
            const float x = xs[i * kKanLanes + lane];
            const float* g = d.p.grid + (static_cast<size_t>(l) * H + i) * d.G;
            const int j = knot_span(g, d.G, x);
            float B[kKanMaxK + 1], dB[kKanMaxK + 1];
            if (j >= 0) span_basis(g, d.G, d.k, j, x, B, dB);
            const float sb = x * sigm(x);
            for (int o = 0; o < H; ++o) {
                const size_t e = ll + static_cast<size_t>(i) * H + o;
                float sp = 0.0f;
                if (j >= 0) {
                    const float* cf = d.p.coef + e * d.C;
#pragma unroll
                    for (int t = 0; t <= kKanMaxK; ++t) {
                        const int c = j - d.k + t;
                        if (t <= d.k && c >= 0 && c < d.C) sp = sp + cf[c] * B[t];
                    }
                }
                const float term = d.p.mask[e] * (d.p.scale_base[e] * sb + d.p.scale_sp[e] * sp);
                ys[o * kKanLanes + lane] = ys[o * kKanLanes + lane] + term;
            }
        }
        for (int o = 0; o < H; ++o) xs[o * kKanLanes + lane] = ys[o * kKanLanes + lane];
    }
}

__global__ __launch_bounds__(kKanLanes) void k_kan_fwd(KanDev d, int n, const float* inputs,
                                                        float* out) {
    extern __shared__ float lds[];
    float* xs = lds;
    float* ys = xs + static_cast<size_t>(d.H) * kKanLanes;
    const int lane = threadIdx.x;
    for (int r = blockIdx.x * kKanLanes + lane; r - lane < n; r += gridDim.x * kKanLanes) {
        if (r >= n) continue;  // the wave stays together; its columns are private
        kan_forward_to(d, inputs + static_cast<size_t>(r) * d.F, d.L, xs, ys, lane);
        for (int p = 0; p < d.P; ++p) {
            float z = 0.0f;
            for (int j = 0; j < d.H; ++j) z = z + xs[j * kKanLanes + lane] * d.p.output_weight[p * d.H + j];
            out[static_cast<size_t>(p) * n + r] = sigm(z + d.p.output_bias[p]);
        }
    }
}

// Gradient accumulators of a workgroup in LDS: dense parameters always, the
// coefficients when coef_in_lds.
struct KanAcc {
    float* w_out;  // [P][H]
    float* b_out;  // [P]
    float* w_in;   // [H][F]
    float* b_in;   // [H]
    float* sb;     // [L][H][H]
    float* ssp;    // [L][H][H]
    float* coef;   // [L][H][H][C] in LDS (coef_in_lds)
    bool coef_in_lds;
};

__device__ __forceinline__ void lds_add(float* a, float v) { atomicAdd(a, v); }
// Global float adds with the hardware atomic (the gradients are plain device
// memory, not fine-grained host-coherent memory).
__device__ __forceinline__ void glb_add(float* a, float v) { unsafeAtomicAdd(a, v); }

__global__ __launch_bounds__(kKanLanes) void k_kan_bwd(KanDev d, int n, const float* inputs,
                                                        const float* gout, ddrs_kan_grads g,
                                                        float* grad_inputs, int coef_in_lds) {
    extern __shared__ float lds[];
    const int H = d.H, F = d.F, P = d.P, L = d.L, C = d.C;
    const int lane = threadIdx.x;
    float* xs = lds;                                  // [H][64]
    float* ys = xs + static_cast<size_t>(H) * kKanLanes;
    float* gh = ys + static_cast<size_t>(H) * kKanLanes;
    float* gx = gh + static_cast<size_t>(H) * kKanLanes;
    KanAcc acc;
    acc.w_out = gx + static_cast<size_t>(H) * kKanLanes;
    acc.b_out = acc.w_out + P * H;
    acc.w_in = acc.b_out + P;
    acc.b_in = acc.w_in + H * F;
    acc.sb = acc.b_in + H;
    acc.ssp = acc.sb + static_cast<size_t>(L) * H * H;
    acc.coef_in_lds = coef_in_lds != 0;
    acc.coef = acc.ssp + static_cast<size_t>(L) * H * H;
    const size_t n_dense = static_cast<size_t>(P) * H + P + static_cast<size_t>(H) * F + H +
                           2 * static_cast<size_t>(L) * H * H;
    const size_t n_lds = n_dense + (acc.coef_in_lds ? static_cast<size_t>(L) * H * H * C : 0);
    for (size_t e = lane; e < n_lds; e += kKanLanes) acc.w_out[e] = 0.0f;
    __syncthreads();

    for (int r = blockIdx.x * kKanLanes + lane; r - lane < n; r += gridDim.x * kKanLanes) {
        if (r >= n) continue;
        const float* in = inputs + static_cast<size_t>(r) * F;
        // output layer and sigmoid
        kan_forward_to(d, in, L, xs, ys, lane);
        for (int j = 0; j < H; ++j) gh[j * kKanLanes + lane] = 0.0f;
        for (int p = 0; p < P; ++p) {
            float z = 0.0f;
            for (int j = 0; j < H; ++j) z = z + xs[j * kKanLanes + lane] * d.p.output_weight[p * H + j];
            const float s = sigm(z + d.p.output_bias[p]);
            const float gz = gout[static_cast<size_t>(p) * n + r] * s * (1.0f - s);
            lds_add(&acc.b_out[p], gz);
            for (int j = 0; j < H; ++j) {
                lds_add(&acc.w_out[p * H + j], gz * xs[j * kKanLanes + lane]);
                gh[j * kKanLanes + lane] = gh[j * kKanLanes + lane] + gz * d.p.output_weight[p * H + j];
            }
        }
        // KAN layers in reverse, each from its recomputed input
        for (int l = L - 1; l >= 0; --l) {
            kan_forward_to(d, in, l, xs, ys, lane);
            const size_t ll = static_cast<size_t>(l) * H * H;
            for (int i = 0; i < H; ++i) {
                const float x = xs[i * kKanLanes + lane];
                const float* gk = d.p.grid + (static_cast<size_t>(l) * H + i) * d.G;
                const int j = knot_span(gk, d.G, x);
                float B[kKanMaxK + 1], dB[kKanMaxK + 1];
                if (j >= 0) span_basis(gk, d.G, d.k, j, x, B, dB);
                const float s = sigm(x);
                const float si = x * s;
                const float dsi = s * (1.0f + x * (1.0f - s));
                float gxi = 0.0f;
                for (int o = 0; o < H; ++o) {
                    const size_t e = ll + static_cast<size_t>(i) * H + o;
                    const float gm = gh[o * kKanLanes + lane] * d.p.mask[e];
                    const float ssp = d.p.scale_sp[e];
                    float sp = 0.0f, dsp = 0.0f;
                    if (j >= 0) {
                        const float* cf = d.p.coef + e * C;
#pragma unroll
                        for (int t = 0; t <= kKanMaxK; ++t) {
                            const int c = j - d.k + t;
                            if (t <= d.k && c >= 0 && c < C) {
                                sp = sp + cf[c] * B[t];
                                dsp = dsp + cf[c] * dB[t];
                                const float v = gm * ssp * B[t];
                                if (acc.coef_in_lds) lds_add(&acc.coef[e * C + c], v);
                                else glb_add(&g.coef[e * C + c], v);
                            }
                        }
                    }
                    lds_add(&acc.sb[e], gm * si);
                    lds_add(&acc.ssp[e], gm * sp);
                    gxi = gxi + gm * (d.p.scale_base[e] * dsi + ssp * dsp);
                }
                gx[i * kKanLanes + lane] = gxi;
            }
            for (int i = 0; i < H; ++i) gh[i * kKanLanes + lane] = gx[i * kKanLanes + lane];
        }
        // input layer
        for (int j = 0; j < H; ++j) {
            const float gj = gh[j * kKanLanes + lane];
            lds_add(&acc.b_in[j], gj);
            for (int f = 0; f < F; ++f) lds_add(&acc.w_in[j * F + f], gj * in[f]);
        }
        if (grad_inputs) {
            for (int f = 0; f < F; ++f) {
                float s = 0.0f;
                for (int j = 0; j < H; ++j) s = s + gh[j * kKanLanes + lane] * d.p.input_weight[j * F + f];
                grad_inputs[static_cast<size_t>(r) * F + f] = s;
            }
        }
    }
    __syncthreads();
    // the block's sums into the global gradients
    auto flush = [&](const float* src, float* dst, size_t cnt) {
        for (size_t e = lane; e < cnt; e += kKanLanes) glb_add(&dst[e], src[e]);
    };
    flush(acc.w_out, g.output_weight, static_cast<size_t>(P) * H);
    flush(acc.b_out, g.output_bias, P);
    flush(acc.w_in, g.input_weight, static_cast<size_t>(H) * F);
    flush(acc.b_in, g.input_bias, H);
    flush(acc.sb, g.scale_base, static_cast<size_t>(L) * H * H);
    flush(acc.ssp, g.scale_sp, static_cast<size_t>(L) * H * H);
    if (acc.coef_in_lds) flush(acc.coef, g.coef, static_cast<size_t>(L) * H * H * C);
}

__global__ void k_kan_zero(float* x, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

ddrs_status kfail(ddrs_status s, const std::string& m) {
    ddrs_internal_set_error(m.c_str());
    return s;
}

ddrs_status check_dims(const ddrs_kan_dims* dims, const ddrs_kan_params* p, KanDev* d) {
    if (!dims || !p) return kfail(DDRS_ERR_INVALID, "kan: NULL dims or params");
    d->F = dims->F;
    d->H = dims->H;
    d->P = dims->P;
    d->L = dims->L;
    d->G = dims->G;
    d->k = dims->k;
    d->C = dims->G - 1 - dims->k;
    if (d->F < 1 || d->H < 1 || d->P < 1 || d->L < 0 || d->k < 0 || d->C < 1)
        return kfail(DDRS_ERR_INVALID, "kan: dimensions must be positive and G > k + 1");
    if (d->H > kKanMaxH || d->P > kKanMaxP || d->L > kKanMaxL || d->k > kKanMaxK)
        return kfail(DDRS_ERR_CAPACITY, "kan: H <= 64, P <= 16, L <= 8 and k <= 5 in this build");
    if (!p->input_weight || !p->input_bias || !p->output_weight || !p->output_bias ||
        (d->L > 0 && (!p->grid || !p->coef || !p->scale_base || !p->scale_sp || !p->mask)))
        return kfail(DDRS_ERR_INVALID, "kan: NULL parameter");
    d->p = *p;
    return DDRS_OK;
}

int kan_grid(int n) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    const int want = (n + kKanLanes - 1) / kKanLanes;
    const int cap = 4 * (ncu > 0 ? ncu : 1);
    return want < cap ? want : cap;
}

}  // namespace
}  // namespace ddrs

using namespace ddrs;

extern "C" {

ddrs_status ddrs_kan_head_fwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, float* d_out, void* stream) {
    KanDev d;
    ddrs_status st = check_dims(dims, params, &d);
    if (st != DDRS_OK) return st;
    if (n < 0) return kfail(DDRS_ERR_INVALID, "kan: n must be >= 0");
    if (n == 0) return DDRS_OK;
    if (!d_inputs || !d_out) return kfail(DDRS_ERR_INVALID, "kan: NULL inputs or outputs");
    const size_t lds = 2 * static_cast<size_t>(d.H) * kKanLanes * sizeof(float);
    hipLaunchKernelGGL(k_kan_fwd, dim3(kan_grid(n)), dim3(kKanLanes), lds,
                       static_cast<hipStream_t>(stream), d, n, d_inputs, d_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kfail(DDRS_ERR_HIP, std::string("kan fwd: ") + hipGetErrorString(e));
    return DDRS_OK;
}

ddrs_status ddrs_kan_head_bwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, const float* d_grad_out,
                              const ddrs_kan_grads* grads, float* d_grad_inputs, void* stream) {
    KanDev d;
    ddrs_status st = check_dims(dims, params, &d);
    if (st != DDRS_OK) return st;
    if (n < 0 || !grads) return kfail(DDRS_ERR_INVALID, "kan: bad n or NULL grads");
    if (!grads->input_weight || !grads->input_bias || !grads->output_weight ||
        !grads->output_bias ||
        (d.L > 0 && (!grads->coef || !grads->scale_base || !grads->scale_sp)))
        return kfail(DDRS_ERR_INVALID, "kan: NULL gradient buffer");
    if (n > 0 && (!d_inputs || !d_grad_out))
        return kfail(DDRS_ERR_INVALID, "kan: NULL inputs or output gradient");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t H = d.H, F = d.F, P = d.P, L = d.L, C = d.C;
    // the gradients are overwritten
    auto zero = [&](float* x, size_t cnt) {
        if (cnt) hipLaunchKernelGGL(k_kan_zero, dim3((cnt + 255) / 256), dim3(256), 0, s, x, cnt);
    };
    zero(grads->input_weight, H * F);
    zero(grads->input_bias, H);
    zero(grads->output_weight, P * H);
    zero(grads->output_bias, P);
    if (L) {
        zero(grads->coef, L * H * H * C);
        zero(grads->scale_base, L * H * H);
        zero(grads->scale_sp, L * H * H);
    }
    if (n > 0) {
        const size_t cols = 4 * H * kKanLanes;
        const size_t dense = P * H + P + H * F + H + 2 * L * H * H;
        const size_t coef = L * H * H * C;
        if ((cols + dense) * sizeof(float) > kKanLds)
            return kfail(DDRS_ERR_CAPACITY, "kan: dimensions exceed the backward's LDS budget");
        const bool coef_in_lds = (cols + dense + coef) * sizeof(float) <= kKanLds;
        const size_t lds = (cols + dense + (coef_in_lds ? coef : 0)) * sizeof(float);
        hipLaunchKernelGGL(k_kan_bwd, dim3(kan_grid(n)), dim3(kKanLanes), lds, s, d, n, d_inputs,
                           d_grad_out, *grads, d_grad_inputs, coef_in_lds ? 1 : 0);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kfail(DDRS_ERR_HIP, std::string("kan bwd: ") + hipGetErrorString(e));
    return DDRS_OK;
}

}  // extern "C"

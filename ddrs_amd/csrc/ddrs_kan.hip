// KAN parameter head on gfx950: forward and reverse mode of
//   h0 = inputs W_in^T + b_in,  h_{l+1} = KanLayer_l(h_l),  out = sigmoid(h_L W_out^T + b_out)
// (reference src/nn/kan_head.rs:187-246). KanLayer follows pykan's KANLayer,
// which the reference's rskan v0.1.3 restates:
//   y[o] = sum_i mask[i,o] (scale_base[i,o] silu(x_i) + scale_sp[i,o] sum_c coef[i,o,c] B_c(x_i))
// with B_c the order-k B-spline basis on the knots of input i (Cox-de Boor).
//
// Kernels:
//  k_kan_prep       the parameters as tables: coefficients transposed to
//                   [l][i][c][o] (o contiguous, padded to HP), mask*scale_base and
//                   mask*scale_sp as [l][i][o], knots with their inverse mean
//                   spacing, the linear layers with o contiguous.
//  k_kan_fwd        one thread per reach; tables staged in LDS (the coefficients
//                   too when they fit), activations in registers.
//  k_kan_bwd_reach  one thread per reach: the forward again, then the reverse
//                   pass through the layers; writes every layer's activation
//                   and its gradient ([layer][unit][reach]) and dL/d inputs.
//  k_kan_bwd_sum    parameter gradients as sums over reaches of outer products
//                   (a GEMM whose inner dimension is the reach), per chunk of
//                   reaches, into per-chunk partials.
//  k_kan_finish     partials summed over chunks in a fixed order and turned into
//                   the parameter gradients, so the backward is deterministic.
// An input lies in one knot span, so only the k+1 basis functions of that span
// are evaluated, by the Cox-de Boor recursion restricted to them (the others
// are exactly zero).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

namespace ddrs {
namespace {

constexpr int kMaxK = 3;        // spline order (a template parameter, 1..3)
constexpr int kMaxH = 32;       // hidden width
constexpr int kMaxP = 16;       // outputs
constexpr int kMaxF = 63;       // input attributes
constexpr int kMaxL = 8;        // KAN layers
constexpr int kMaxC1 = 64;      // basis functions per input + 1
constexpr int kNT = 256;        // threads per workgroup
constexpr int kSub = 128;       // reaches per LDS tile of the gradient sums
constexpr size_t kLdsBudget = 65536;

__host__ __device__ constexpr int align4(int x) { return (x + 3) & ~3; }

// Table layout in floats; the LDS copy is the prefix [0, coef) or all of it.
struct KanLayout {
    int F, H, P, L, G, k, C, HP;
    int wsb, wsp, grid, ginv, win, bin, wout, bout, coef, total;
    void make(int F_, int H_, int P_, int L_, int G_, int k_, int HP_) {
        F = F_; H = H_; P = P_; L = L_; G = G_; k = k_; C = G_ - 1 - k_; HP = HP_;
        wsb = 0;
        wsp = wsb + L * H * HP;
        grid = wsp + L * H * HP;
        ginv = align4(grid + L * H * G);
        win = align4(ginv + L * H);
        bin = win + F * HP;
        wout = bin + HP;
        bout = wout + P * HP;
        coef = align4(bout + P);
        total = coef + L * H * C * HP;
    }
};

struct KanReachArgs {
    KanLayout lay;
    int n;
    const float* tab;      // prepared tables (global)
    const float* inputs;   // [n][F]
    float* out;            // [P][n] (forward)
    // reverse pass
    const float* gout;     // [P][n]
    float* acts;           // [L+1][H][n]
    float* dacts;          // [L+1][H][n]
    float* gz;             // [P][n]
    float* grad_inputs;    // [n][F] or null
};

// Sigmoid with the hardware exponential and reciprocal (about 2 ulp).
__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// Knot span of x: j with g[j] <= x < g[j+1], or -1 outside [g[0], g[G-1]).
// Starts from the span of a uniform grid and corrects it, so the result is
// exact for any strictly ascending knots.
__device__ __forceinline__ int knot_span(const float* g, int G, float ginv, float x) {
    if (!(x >= g[0] && x < g[G - 1])) return -1;
    int j = static_cast<int>((x - g[0]) * ginv);
    j = min(max(j, 0), G - 2);
    while (j < G - 2 && g[j + 1] <= x) ++j;
    while (j > 0 && g[j] > x) --j;
    return j;
}

// The K+1 basis functions that can be non-zero at x in span j, B[t] = B_{j-K+t}
// (and their x-derivatives when D): the Cox-de Boor recursion with its
// divisions taken as products with hardware reciprocals of the knot
// differences. At order kk the basis has G-1-kk functions (c <= G-2-kk);
// terms outside are zero. All zero when j < 0 (x outside the knots). The knots
// a term reads lie in the window g[j-K .. j+K+1], loaded once (clamped reads).
template <int K, bool D>
__device__ __forceinline__ void span_basis(const float* g, int G, int j, float x,
                                           float (&B)[K + 1], float (&dB)[K + 1]) {
    const int j0 = max(j, 0);
    float w[2 * K + 2];
#pragma unroll
    for (int m = 0; m < 2 * K + 2; ++m) w[m] = g[min(max(j0 - K + m, 0), G - 1)];
#pragma unroll
    for (int t = 0; t <= K; ++t) {
        B[t] = (t == K && j >= 0) ? 1.0f : 0.0f;
        dB[t] = 0.0f;
    }
#pragma unroll
    for (int kk = 1; kk <= K; ++kk) {
        float r[K + 2];
#pragma unroll
        for (int t = K - kk; t <= K + 1; ++t) r[t] = __builtin_amdgcn_rcpf(w[t + kk] - w[t]);
#pragma unroll
        for (int t = K - kk; t <= K; ++t) {
            const int c = j0 - K + t;
            const bool ok = c >= 0 && c <= G - 2 - kk;
            const float b0 = B[t];
            const float b1 = (t < K) ? B[t + 1] : 0.0f;
            const float a = (x - w[t]) * r[t];
            const float cc = (w[t + kk + 1] - x) * r[t + 1];
            const float v = a * b0 + cc * b1;
            if (D) {
                const float db0 = dB[t];
                const float db1 = (t < K) ? dB[t + 1] : 0.0f;
                const float dv = ((b0 * r[t] + a * db0) - b1 * r[t + 1]) + cc * db1;
                dB[t] = ok ? dv : 0.0f;
            }
            B[t] = ok ? v : 0.0f;
        }
    }
}

// acc[o] = sum_t coefT[row][j-K+t][o] * W[t] (W = B or dB; zero when j < 0).
template <int HP, int K>
__device__ __forceinline__ void spline_rows(const float* coefT, int row, int C, int j,
                                            const float (&W)[K + 1], float (&acc)[HP]) {
#pragma unroll
    for (int o = 0; o < HP; ++o) acc[o] = 0.0f;
#pragma unroll
    for (int t = 0; t <= K; ++t) {
        // outside [0, C) the weight is exactly zero: clamp the row index
        const int c = min(max(j - K + t, 0), C - 1);
        const float4* r4 = reinterpret_cast<const float4*>(coefT + (row * C + c) * HP);
#pragma unroll
        for (int q = 0; q < HP / 4; ++q) {
            const float4 w = r4[q];
            acc[4 * q + 0] = fmaf(w.x, W[t], acc[4 * q + 0]);
            acc[4 * q + 1] = fmaf(w.y, W[t], acc[4 * q + 1]);
            acc[4 * q + 2] = fmaf(w.z, W[t], acc[4 * q + 2]);
            acc[4 * q + 3] = fmaf(w.w, W[t], acc[4 * q + 3]);
        }
    }
}

// The first nfl floats of the tables into LDS (nfl a multiple of 4).
__device__ __forceinline__ void stage_tables(const float* tab, float* lds, int nfl) {
    const float4* s = reinterpret_cast<const float4*>(tab);
    float4* d = reinterpret_cast<float4*>(lds);
    for (int e = threadIdx.x; e < nfl / 4; e += blockDim.x) d[e] = s[e];
}

template <int HP>
__device__ __forceinline__ void input_layer(const KanLayout& Y, const float* T, const float* in,
                                            float (&x)[HP]) {
#pragma unroll
    for (int o = 0; o < HP; ++o) x[o] = 0.0f;
    for (int f = 0; f < Y.F; ++f) {
        const float v = in[f];
        const float4* w4 = reinterpret_cast<const float4*>(T + Y.win + f * HP);
#pragma unroll
        for (int q = 0; q < HP / 4; ++q) {
            const float4 w = w4[q];
            x[4 * q + 0] = fmaf(v, w.x, x[4 * q + 0]);
            x[4 * q + 1] = fmaf(v, w.y, x[4 * q + 1]);
            x[4 * q + 2] = fmaf(v, w.z, x[4 * q + 2]);
            x[4 * q + 3] = fmaf(v, w.w, x[4 * q + 3]);
        }
    }
#pragma unroll
    for (int o = 0; o < HP; ++o) x[o] = x[o] + T[Y.bin + o];
}

// KAN layers of one reach: x (layer 0 input) -> the last activation, in place.
// With acts != null every layer's input and the last activation are stored.
template <int HP, int K>
__device__ __forceinline__ void kan_layers(const KanLayout& Y, const float* T, const float* coefT,
                                           float (&x)[HP], float* acts, size_t n, size_t r) {
    for (int l = 0; l < Y.L; ++l) {
        if (acts) {
#pragma unroll
            for (int i = 0; i < HP; ++i)
                if (i < Y.H) acts[(static_cast<size_t>(l) * Y.H + i) * n + r] = x[i];
        }
        float y[HP];
#pragma unroll
        for (int o = 0; o < HP; ++o) y[o] = 0.0f;
        // rolled: the body stays small enough for the instruction cache
#pragma unroll 1
        for (int i = 0; i < Y.H; ++i) {
            {
                const int row = l * Y.H + i;
                const float xi = x[i];
                const float* g = T + Y.grid + row * Y.G;
                const int j = knot_span(g, Y.G, T[Y.ginv + row], xi);
                float B[K + 1], dB[K + 1];
                span_basis<K, false>(g, Y.G, j, xi, B, dB);
                float sp[HP];
                spline_rows<HP, K>(coefT, row, Y.C, j, B, sp);
                const float s = xi * sigm(xi);
                const float4* sb4 = reinterpret_cast<const float4*>(T + Y.wsb + row * HP);
                const float4* ss4 = reinterpret_cast<const float4*>(T + Y.wsp + row * HP);
#pragma unroll
                for (int q = 0; q < HP / 4; ++q) {
                    const float4 a = sb4[q], b = ss4[q];
                    y[4 * q + 0] = fmaf(b.x, sp[4 * q + 0], fmaf(a.x, s, y[4 * q + 0]));
                    y[4 * q + 1] = fmaf(b.y, sp[4 * q + 1], fmaf(a.y, s, y[4 * q + 1]));
                    y[4 * q + 2] = fmaf(b.z, sp[4 * q + 2], fmaf(a.z, s, y[4 * q + 2]));
                    y[4 * q + 3] = fmaf(b.w, sp[4 * q + 3], fmaf(a.w, s, y[4 * q + 3]));
                }
            }
        }
#pragma unroll
        for (int o = 0; o < HP; ++o) x[o] = y[o];
    }
    if (acts) {
#pragma unroll
        for (int i = 0; i < HP; ++i)
            if (i < Y.H) acts[(static_cast<size_t>(Y.L) * Y.H + i) * n + r] = x[i];
    }
}

// CL: the coefficient table is staged in LDS (else read from global memory).
template <int HP, int K, bool CL>
__global__ __launch_bounds__(kNT) void k_kan_fwd(KanReachArgs a) {
    extern __shared__ float4 lds4[];
    float* T = reinterpret_cast<float*>(lds4);
    const KanLayout Y = a.lay;
    stage_tables(a.tab, T, CL ? Y.total : Y.coef);
    __syncthreads();
    const float* coefT = CL ? T + Y.coef : a.tab + Y.coef;
    const size_t n = a.n;
    for (size_t r = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x; r < n;
         r += static_cast<size_t>(gridDim.x) * kNT) {
        float x[HP];
        input_layer<HP>(Y, T, a.inputs + r * Y.F, x);
        kan_layers<HP, K>(Y, T, coefT, x, nullptr, n, r);
        for (int p = 0; p < Y.P; ++p) {
            float z = 0.0f;
#pragma unroll
            for (int j = 0; j < HP; ++j) z = fmaf(x[j], T[Y.wout + p * HP + j], z);
            a.out[static_cast<size_t>(p) * n + r] = sigm(z + T[Y.bout + p]);
        }
    }
}

template <int HP, int K, bool CL>
__global__ __launch_bounds__(kNT) void k_kan_bwd_reach(KanReachArgs a) {
    extern __shared__ float4 lds4[];
    float* T = reinterpret_cast<float*>(lds4);
    const KanLayout Y = a.lay;
    stage_tables(a.tab, T, CL ? Y.total : Y.coef);
    __syncthreads();
    const float* coefT = CL ? T + Y.coef : a.tab + Y.coef;
    const size_t n = a.n;
    for (size_t r = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x; r < n;
         r += static_cast<size_t>(gridDim.x) * kNT) {
        float x[HP];
        input_layer<HP>(Y, T, a.inputs + r * Y.F, x);
        kan_layers<HP, K>(Y, T, coefT, x, a.acts, n, r);
        // output layer and sigmoid
        float gh[HP];
#pragma unroll
        for (int j = 0; j < HP; ++j) gh[j] = 0.0f;
        for (int p = 0; p < Y.P; ++p) {
            float z = 0.0f;
#pragma unroll
            for (int j = 0; j < HP; ++j) z = fmaf(x[j], T[Y.wout + p * HP + j], z);
            const float s = sigm(z + T[Y.bout + p]);
            const float gzp = a.gout[static_cast<size_t>(p) * n + r] * s * (1.0f - s);
            a.gz[static_cast<size_t>(p) * n + r] = gzp;
#pragma unroll
            for (int j = 0; j < HP; ++j) gh[j] = fmaf(gzp, T[Y.wout + p * HP + j], gh[j]);
        }
        // KAN layers in reverse, each from its stored input; gh = dL/d(layer output)
        for (int l = Y.L; l >= 0; --l) {
#pragma unroll
            for (int j = 0; j < HP; ++j)
                if (j < Y.H) a.dacts[(static_cast<size_t>(l) * Y.H + j) * n + r] = gh[j];
            if (l == 0) break;
            const int lp = l - 1;
            float gx[HP];
#pragma unroll
            for (int i = 0; i < HP; ++i) gx[i] = 0.0f;
#pragma unroll 1
            for (int i = 0; i < Y.H; ++i) {
                {
                    const int row = lp * Y.H + i;
                    const float xi = a.acts[(static_cast<size_t>(lp) * Y.H + i) * n + r];
                    const float* g = T + Y.grid + row * Y.G;
                    const int j = knot_span(g, Y.G, T[Y.ginv + row], xi);
                    float B[K + 1], dB[K + 1];
                    span_basis<K, true>(g, Y.G, j, xi, B, dB);
                    float dsp[HP];
                    spline_rows<HP, K>(coefT, row, Y.C, j, dB, dsp);
                    const float s = sigm(xi);
                    const float ds = s * (1.0f + xi * (1.0f - s));
                    const float4* sb4 = reinterpret_cast<const float4*>(T + Y.wsb + row * HP);
                    const float4* ss4 = reinterpret_cast<const float4*>(T + Y.wsp + row * HP);
                    float acc = 0.0f;
#pragma unroll
                    for (int q = 0; q < HP / 4; ++q) {
                        const float4 u = sb4[q], v = ss4[q];
                        acc = fmaf(gh[4 * q + 0], fmaf(v.x, dsp[4 * q + 0], u.x * ds), acc);
                        acc = fmaf(gh[4 * q + 1], fmaf(v.y, dsp[4 * q + 1], u.y * ds), acc);
                        acc = fmaf(gh[4 * q + 2], fmaf(v.z, dsp[4 * q + 2], u.z * ds), acc);
                        acc = fmaf(gh[4 * q + 3], fmaf(v.w, dsp[4 * q + 3], u.w * ds), acc);
                    }
                    gx[i] = acc;
                }
            }
#pragma unroll
            for (int j = 0; j < HP; ++j) gh[j] = gx[j];
        }
        if (a.grad_inputs) {
            for (int f = 0; f < Y.F; ++f) {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < HP; ++j) s = fmaf(gh[j], T[Y.win + f * HP + j], s);
                a.grad_inputs[r * Y.F + f] = s;
            }
        }
    }
}

// Tables from the caller's parameters.
__global__ void k_kan_prep(KanLayout Y, ddrs_kan_params p, float* tab) {
    const int HP = Y.HP, H = Y.H, C = Y.C;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < Y.total; e += gridDim.x * blockDim.x) {
        float v = 0.0f;
        if (e < Y.wsp) {                       // mask * scale_base, [l][i][o]
            const int o = e % HP, li = e / HP;
            if (o < H) v = p.mask[li * H + o] * p.scale_base[li * H + o];
        } else if (e < Y.grid) {               // mask * scale_sp
            const int q = e - Y.wsp, o = q % HP, li = q / HP;
            if (o < H) v = p.mask[li * H + o] * p.scale_sp[li * H + o];
        } else if (e < Y.ginv) {
            const int q = e - Y.grid;
            if (q < Y.L * H * Y.G) v = p.grid[q];
        } else if (e < Y.win) {
            const int q = e - Y.ginv;
            if (q < Y.L * H) {
                const float* g = p.grid + q * Y.G;
                v = static_cast<float>(Y.G - 1) / (g[Y.G - 1] - g[0]);
            }
        } else if (e < Y.bin) {                // W_in^T, [f][o]
            const int q = e - Y.win, o = q % HP, f = q / HP;
            if (o < H) v = p.input_weight[o * Y.F + f];
        } else if (e < Y.wout) {
            const int o = e - Y.bin;
            if (o < H) v = p.input_bias[o];
        } else if (e < Y.bout) {               // W_out, [p][j]
            const int q = e - Y.wout, j = q % HP, pp = q / HP;
            if (j < H) v = p.output_weight[pp * H + j];
        } else if (e < Y.coef) {
            const int q = e - Y.bout;
            if (q < Y.P) v = p.output_bias[q];
        } else {                               // coef^T, [l][i][c][o]
            const int q = e - Y.coef, o = q % HP, c = (q / HP) % C, li = q / (HP * C);
            if (o < H) v = p.coef[(static_cast<size_t>(li) * H + o) * C + c];
        }
        tab[e] = v;
    }
}

// Gradient sums, out[u][v] = sum_r U[r][u] V[r][v] over a chunk of reaches:
// item (l, i) < L*H has U = (B_0..B_{C-1}, silu) of layer l's input i and
// V = dL/d(layer l's outputs); item L*H the output layer (U = [h_L, 1],
// V = dL/dz); item L*H+1 the input layer (U = [inputs, 1], V = dL/dh0).
// One workgroup per (chunk, item).
struct KanSumArgs {
    KanLayout lay;
    int n, chunk, nch, items, nblocks, E;
    const float* tab;
    const float* inputs;
    const float* acts;
    const float* dacts;
    const float* gz;
    float* partial;        // [E][nch]
};

template <int K>
__global__ __launch_bounds__(kNT) void k_kan_bwd_sum(KanSumArgs a) {
    extern __shared__ float4 lds4[];
    float* sm = reinterpret_cast<float*>(lds4);
    const KanLayout Y = a.lay;
    // XCD-aware order: workgroups b, b+8, b+16, ... share an XCD and take
    // consecutive (chunk, item) pairs, so the items of a chunk share its L2.
    const int b = blockIdx.x;
    const int per = a.nblocks / 8;
    const int logical = (b % 8) * per + b / 8;
    const int ch = logical / a.items, it = logical % a.items;
    if (ch >= a.nch) return;
    const int H = Y.H, L = Y.L, C = Y.C;
    int UW, VW, base, kind, l = 0, i = 0;  // kind: 0 KAN, 1 output layer, 2 input layer
    if (it < L * H) {
        kind = 0; l = it / H; i = it % H;
        UW = C + 1; VW = H;
        base = (l * H + i) * (C + 1) * H;
    } else if (it == L * H) {
        kind = 1; UW = H + 1; VW = Y.P;
        base = L * H * (C + 1) * H;
    } else {
        kind = 2; UW = Y.F + 1; VW = H;
        base = L * H * (C + 1) * H + Y.P * (H + 1);
    }
    const int UWP = align4(UW), VWP = align4(VW);
    float* Us = sm;                   // [kSub][UWP]
    float* Vs = sm + kSub * UWP;      // [kSub][VWP]
    const int nbu = UWP / 4, nbv = VWP / 4, nblk = nbu * nbv;
    const int nks = kNT / nblk;
    const int tid = threadIdx.x;
    const int blk = tid % nblk, ks = tid / nblk;
    const int ub = blk % nbu, vb = blk / nbu;
    float acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0f;
    const size_t n = a.n;
    const size_t r_begin = static_cast<size_t>(ch) * a.chunk;
    const size_t r_end = min(n, r_begin + static_cast<size_t>(a.chunk));
    const float* g = a.tab + Y.grid + (l * H + i) * Y.G;
    const float ginv = a.tab[Y.ginv + l * H + i];
    for (size_t r0 = r_begin; r0 < r_end; r0 += kSub) {
        __syncthreads();
        for (int e = tid; e < kSub * VWP; e += kNT) {
            const int t = e % kSub, o = e / kSub;
            const size_t r = r0 + t;
            float v = 0.0f;
            if (r < r_end && o < VW) {
                if (kind == 0) v = a.dacts[(static_cast<size_t>(l + 1) * H + o) * n + r];
                else if (kind == 1) v = a.gz[static_cast<size_t>(o) * n + r];
                else v = a.dacts[static_cast<size_t>(o) * n + r];
            }
            Vs[t * VWP + o] = v;
        }
        if (kind == 0) {
            for (int t = tid; t < kSub; t += kNT) {
                const size_t r = r0 + t;
                float* u = Us + t * UWP;
                for (int c = 0; c < UWP; ++c) u[c] = 0.0f;
                if (r < r_end) {
                    const float x = a.acts[(static_cast<size_t>(l) * H + i) * n + r];
                    const int j = knot_span(g, Y.G, ginv, x);
                    float B[K + 1], dB[K + 1];
                    span_basis<K, false>(g, Y.G, j, x, B, dB);
                    if (j >= 0) {
#pragma unroll
                        for (int tt = 0; tt <= K; ++tt) {
                            const int c = j - K + tt;
                            if (c >= 0 && c < C) u[c] = B[tt];
                        }
                    }
                    u[C] = x * sigm(x);
                }
            }
        } else {
            for (int e = tid; e < kSub * UWP; e += kNT) {
                const int t = e % kSub, c = e / kSub;
                const size_t r = r0 + t;
                float v = 0.0f;
                if (r < r_end && c < UW) {
                    if (c == UW - 1) v = 1.0f;
                    else if (kind == 1) v = a.acts[(static_cast<size_t>(L) * H + c) * n + r];
                    else v = a.inputs[r * Y.F + c];
                }
                Us[t * UWP + c] = v;
            }
        }
        __syncthreads();
        if (ks < nks) {
            const float4* U4 = reinterpret_cast<const float4*>(Us);
            const float4* V4 = reinterpret_cast<const float4*>(Vs);
            for (int t = ks; t < kSub; t += nks) {
                const float4 u = U4[t * nbu + ub];
                const float4 v = V4[t * nbv + vb];
                const float uu[4] = {u.x, u.y, u.z, u.w};
                const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[p][q] = fmaf(uu[p], vv[q], acc[p][q]);
            }
        }
    }
    // sum the thread slices of the reach dimension, in order
    __syncthreads();
    float* red = sm;  // [nks][nblk][16]
    if (ks < nks) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[(ks * nblk + blk) * 16 + p * 4 + q] = acc[p][q];
    }
    __syncthreads();
    float* dst = a.partial + static_cast<size_t>(base) * a.nch + ch;
    for (int e = tid; e < nblk * 16; e += kNT) {
        const int bb = e / 16, pq = e % 16;
        float s = 0.0f;
        for (int k2 = 0; k2 < nks; ++k2) s += red[(k2 * nblk + bb) * 16 + pq];
        const int uu = (bb % nbu) * 4 + pq / 4, vv = (bb / nbu) * 4 + pq % 4;
        if (uu < UW && vv < VW) {
            const int e = kind == 0 ? uu * H + vv : vv * UW + uu;  // [c][o] or [out][in, bias]
            dst[static_cast<size_t>(e) * a.nch] = s;
        }
    }
}

// Partials summed over chunks in chunk order: S[e] = sum_ch partial[e][ch].
__global__ void k_kan_sum_chunks(const float* partial, int nch, int E, float* S) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float* p = partial + static_cast<size_t>(e) * nch;
    float s = 0.0f;
    for (int c = 0; c < nch; ++c) s += p[c];
    S[e] = s;
}

// The parameter gradients from the sums: coef = mask*ssp*S_c,
// scale_sp = mask*sum_c coef*S_c, scale_base = mask*S_silu, and the linear
// layers' weights and biases.
__global__ void k_kan_finish(KanLayout Y, ddrs_kan_params p, ddrs_kan_grads g,
                             const float* tab, const float* S) {
    const int H = Y.H, C = Y.C, L = Y.L;
    const int n_kan = L * H * H;
    const int n_out = Y.P * (H + 1);
    const int n_in = H * (Y.F + 1);
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n_kan) {
        const int o = u % H, li = u / H;         // li = l*H + i
        const float* Sk = S + li * (C + 1) * H;  // [c][o]
        const float wsp = tab[Y.wsp + li * Y.HP + o];
        float gsp = 0.0f;
        for (int c = 0; c < C; ++c) {
            const float v = Sk[c * H + o];
            g.coef[static_cast<size_t>(u) * C + c] = wsp * v;
            gsp = fmaf(p.coef[static_cast<size_t>(u) * C + c], v, gsp);
        }
        g.scale_sp[u] = p.mask[u] * gsp;
        g.scale_base[u] = p.mask[u] * Sk[C * H + o];
    } else if (u < n_kan + n_out) {
        const int q = u - n_kan, pp = q / (H + 1), j = q % (H + 1);
        const float s = S[L * H * (C + 1) * H + q];
        if (j < H) g.output_weight[pp * H + j] = s;
        else g.output_bias[pp] = s;
    } else if (u < n_kan + n_out + n_in) {
        const int q = u - n_kan - n_out, j = q / (Y.F + 1), f = q % (Y.F + 1);
        const float s = S[L * H * (C + 1) * H + n_out + q];
        if (f < Y.F) g.input_weight[j * Y.F + f] = s;
        else g.input_bias[j] = s;
    }
}

__global__ void k_kan_zero(float* x, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

ddrs_status kfail(ddrs_status s, const std::string& m) {
    ddrs_internal_set_error(m.c_str());
    return s;
}

int hp_of(int H) { return H <= 8 ? 8 : H <= 16 ? 16 : H <= 24 ? 24 : 32; }

ddrs_status check_dims(const ddrs_kan_dims* d, KanLayout* Y) {
    if (!d) return kfail(DDRS_ERR_INVALID, "kan: NULL dims");
    const int C = d->G - 1 - d->k;
    if (d->F < 1 || d->H < 1 || d->P < 1 || d->L < 0 || d->k < 0 || C < 1)
        return kfail(DDRS_ERR_INVALID, "kan: dimensions must be positive and G > k + 1");
    if (d->H > kMaxH || d->P > kMaxP || d->F > kMaxF || d->L > kMaxL || d->k > kMaxK ||
        d->k < 1 ||
        C + 1 > kMaxC1)
        return kfail(DDRS_ERR_CAPACITY,
                     "kan: H <= 32, P <= 16, F <= 63, L <= 8, 1 <= k <= 3 and G - k <= 64 in this build");
    Y->make(d->F, d->H, d->P, d->L, d->G, d->k, hp_of(d->H));
    return DDRS_OK;
}

ddrs_status check_params(const KanLayout& Y, const ddrs_kan_params* p) {
    if (!p) return kfail(DDRS_ERR_INVALID, "kan: NULL params");
    if (!p->input_weight || !p->input_bias || !p->output_weight || !p->output_bias ||
        (Y.L > 0 && (!p->grid || !p->coef || !p->scale_base || !p->scale_sp || !p->mask)))
        return kfail(DDRS_ERR_INVALID, "kan: NULL parameter");
    return DDRS_OK;
}

size_t up256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Workspace: tables | acts | dacts | gz | partials (the forward needs the tables).
struct KanWs {
    size_t tab, acts, dacts, gz, partial, sums, fwd_bytes, bwd_bytes;
    int nch, chunk, E;
};

KanWs ws_layout(const KanLayout& Y, int n) {
    KanWs w{};
    const size_t N = static_cast<size_t>(n > 0 ? n : 0);
    w.tab = 0;
    w.acts = up256(static_cast<size_t>(Y.total) * 4);
    w.fwd_bytes = w.acts;
    const size_t la = static_cast<size_t>(Y.L + 1) * Y.H * N * 4;
    w.dacts = w.acts + up256(la);
    w.gz = w.dacts + up256(la);
    w.partial = w.gz + up256(static_cast<size_t>(Y.P) * N * 4);
    // chunks of at least 2048 reaches, at most 128 of them
    const int want = (n + 2047) / 2048;
    w.nch = want < 1 ? 1 : (want > 128 ? 128 : want);
    w.chunk = (n + w.nch - 1) / w.nch;
    if (w.chunk < 1) w.chunk = 1;
    w.E = Y.L * Y.H * (Y.C + 1) * Y.H + Y.P * (Y.H + 1) + Y.H * (Y.F + 1);
    w.sums = w.partial + up256(static_cast<size_t>(w.nch) * w.E * 4);
    w.bwd_bytes = w.sums + up256(static_cast<size_t>(w.E) * 4);
    return w;
}

int n_cus() {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
        ncu = 256;
    return ncu;
}

template <int HP, int K, bool CL>
void launch_reach(bool bwd, const KanReachArgs& a, int grid, size_t lds, hipStream_t s) {
    if (bwd) hipLaunchKernelGGL((k_kan_bwd_reach<HP, K, CL>), dim3(grid), dim3(kNT), lds, s, a);
    else hipLaunchKernelGGL((k_kan_fwd<HP, K, CL>), dim3(grid), dim3(kNT), lds, s, a);
}

template <int K, bool CL>
void launch_reach_hp(bool bwd, const KanReachArgs& a, int grid, size_t lds, hipStream_t s) {
    switch (a.lay.HP) {
        case 8: launch_reach<8, K, CL>(bwd, a, grid, lds, s); break;
        case 16: launch_reach<16, K, CL>(bwd, a, grid, lds, s); break;
        case 24: launch_reach<24, K, CL>(bwd, a, grid, lds, s); break;
        default: launch_reach<32, K, CL>(bwd, a, grid, lds, s); break;
    }
}

template <bool CL>
void launch_reach_k(bool bwd, const KanReachArgs& a, int grid, size_t lds, hipStream_t s) {
    switch (a.lay.k) {
        case 1: launch_reach_hp<1, CL>(bwd, a, grid, lds, s); break;
        case 2: launch_reach_hp<2, CL>(bwd, a, grid, lds, s); break;
        default: launch_reach_hp<3, CL>(bwd, a, grid, lds, s); break;
    }
}

// Checks, the tables into the workspace, and the reach kernels' launch.
struct KanPlan {
    KanLayout Y;
    KanWs W;
    KanReachArgs a;
    int grid;
    size_t lds;
    bool coef_lds;
};

ddrs_status prepare(const ddrs_kan_dims* dims, const ddrs_kan_params* params, int32_t n,
                    const float* d_inputs, void* d_ws, size_t ws_bytes, bool bwd, KanPlan* pl,
                    hipStream_t s) {
    ddrs_status st = check_dims(dims, &pl->Y);
    if (st != DDRS_OK) return st;
    st = check_params(pl->Y, params);
    if (st != DDRS_OK) return st;
    if (n < 0) return kfail(DDRS_ERR_INVALID, "kan: n must be >= 0");
    pl->W = ws_layout(pl->Y, n);
    if (!d_ws || ws_bytes < (bwd ? pl->W.bwd_bytes : pl->W.fwd_bytes))
        return kfail(DDRS_ERR_INVALID,
                     "kan: workspace missing or smaller than ddrs_kan_workspace_bytes");
    if (n > 0 && !d_inputs) return kfail(DDRS_ERR_INVALID, "kan: NULL inputs");
    const KanLayout& Y = pl->Y;
    float* tab = reinterpret_cast<float*>(static_cast<char*>(d_ws) + pl->W.tab);
    hipLaunchKernelGGL(k_kan_prep, dim3((Y.total + 255) / 256), dim3(256), 0, s, Y, *params, tab);
    pl->a = KanReachArgs{};
    pl->a.lay = Y;
    pl->a.n = n;
    pl->a.tab = tab;
    pl->a.inputs = d_inputs;
    const size_t full = static_cast<size_t>(Y.total) * 4;
    pl->coef_lds = full <= kLdsBudget;
    pl->lds = pl->coef_lds ? full : static_cast<size_t>(Y.coef) * 4;
    if (pl->lds > kLdsBudget) return kfail(DDRS_ERR_CAPACITY, "kan: tables exceed LDS");
    const int want = (n + kNT - 1) / kNT;
    const int cap = 4 * n_cus();
    pl->grid = want < cap ? want : cap;
    return DDRS_OK;
}

void launch_reach_plan(bool bwd, const KanPlan& pl, hipStream_t s) {
    if (pl.coef_lds) launch_reach_k<true>(bwd, pl.a, pl.grid, pl.lds, s);
    else launch_reach_k<false>(bwd, pl.a, pl.grid, pl.lds, s);
}

}  // namespace
}  // namespace ddrs

using namespace ddrs;

extern "C" {

ddrs_status ddrs_kan_workspace_bytes(const ddrs_kan_dims* dims, int32_t n, size_t* fwd_bytes,
                                     size_t* bwd_bytes) {
    KanLayout Y;
    ddrs_status st = check_dims(dims, &Y);
    if (st != DDRS_OK) return st;
    if (n < 0) return kfail(DDRS_ERR_INVALID, "kan: n must be >= 0");
    const KanWs W = ws_layout(Y, n);
    if (fwd_bytes) *fwd_bytes = W.fwd_bytes;
    if (bwd_bytes) *bwd_bytes = W.bwd_bytes;
    return DDRS_OK;
}

ddrs_status ddrs_kan_head_fwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, float* d_out, void* d_workspace,
                              size_t workspace_bytes, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    KanPlan pl;
    ddrs_status st0 = check_dims(dims, &pl.Y);
    if (st0 != DDRS_OK) return st0;
    if (n > 0 && !d_out) return kfail(DDRS_ERR_INVALID, "kan: NULL outputs");
    ddrs_status st =
        prepare(dims, params, n, d_inputs, d_workspace, workspace_bytes, false, &pl, s);
    if (st != DDRS_OK) return st;
    if (n > 0) {
        pl.a.out = d_out;
        launch_reach_plan(false, pl, s);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kfail(DDRS_ERR_HIP, std::string("kan fwd: ") + hipGetErrorString(e));
    return DDRS_OK;
}

ddrs_status ddrs_kan_head_bwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, const float* d_grad_out,
                              const ddrs_kan_grads* grads, float* d_grad_inputs,
                              void* d_workspace, size_t workspace_bytes, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    KanPlan pl;
    ddrs_status st0 = check_dims(dims, &pl.Y);
    if (st0 != DDRS_OK) return st0;
    if (!grads || !grads->input_weight || !grads->input_bias || !grads->output_weight ||
        !grads->output_bias)
        return kfail(DDRS_ERR_INVALID, "kan: NULL gradient buffer");
    if (dims->L > 0 && (!grads->coef || !grads->scale_base || !grads->scale_sp))
        return kfail(DDRS_ERR_INVALID, "kan: NULL gradient buffer");
    if (n > 0 && !d_grad_out) return kfail(DDRS_ERR_INVALID, "kan: NULL output gradient");
    ddrs_status st = prepare(dims, params, n, d_inputs, d_workspace, workspace_bytes, true, &pl, s);
    if (st != DDRS_OK) return st;
    const KanLayout& Y = pl.Y;
    const KanWs& W = pl.W;
    char* ws = static_cast<char*>(d_workspace);
    if (n == 0) {
        auto zero = [&](float* x, size_t cnt) {
            if (cnt) hipLaunchKernelGGL(k_kan_zero, dim3((cnt + 255) / 256), dim3(256), 0, s, x, cnt);
        };
        const size_t H = Y.H, F = Y.F, P = Y.P, L = Y.L, C = Y.C;
        zero(grads->input_weight, H * F);
        zero(grads->input_bias, H);
        zero(grads->output_weight, P * H);
        zero(grads->output_bias, P);
        if (L) {
            zero(grads->coef, L * H * H * C);
            zero(grads->scale_base, L * H * H);
            zero(grads->scale_sp, L * H * H);
        }
    } else {
        pl.a.gout = d_grad_out;
        pl.a.acts = reinterpret_cast<float*>(ws + W.acts);
        pl.a.dacts = reinterpret_cast<float*>(ws + W.dacts);
        pl.a.gz = reinterpret_cast<float*>(ws + W.gz);
        pl.a.grad_inputs = d_grad_inputs;
        launch_reach_plan(true, pl, s);

        KanSumArgs b{};
        b.lay = Y;
        b.n = n;
        b.chunk = W.chunk;
        b.nch = W.nch;
        b.items = Y.L * Y.H + 2;
        b.nblocks = (W.nch * b.items + 7) / 8 * 8;
        b.E = W.E;
        b.tab = pl.a.tab;
        b.inputs = d_inputs;
        b.acts = pl.a.acts;
        b.dacts = pl.a.dacts;
        b.gz = pl.a.gz;
        b.partial = reinterpret_cast<float*>(ws + W.partial);
        // the widest U and V rows of any item, or the reduction buffer
        const int uwp = align4(std::max(std::max(Y.C + 1, Y.H + 1), Y.F + 1));
        const int vwp = align4(std::max(Y.H, Y.P));
        const size_t lds_sum = std::max<size_t>(static_cast<size_t>(kSub) * (uwp + vwp),
                                                static_cast<size_t>(kNT) * 16) * 4;
        switch (Y.k) {
            case 1: hipLaunchKernelGGL(k_kan_bwd_sum<1>, dim3(b.nblocks), dim3(kNT), lds_sum, s, b); break;
            case 2: hipLaunchKernelGGL(k_kan_bwd_sum<2>, dim3(b.nblocks), dim3(kNT), lds_sum, s, b); break;
            default: hipLaunchKernelGGL(k_kan_bwd_sum<3>, dim3(b.nblocks), dim3(kNT), lds_sum, s, b); break;
        }
        float* sums = reinterpret_cast<float*>(ws + W.sums);
        hipLaunchKernelGGL(k_kan_sum_chunks, dim3((W.E + 255) / 256), dim3(256), 0, s, b.partial,
                           W.nch, W.E, sums);
        const int units = Y.L * Y.H * Y.H + Y.P * (Y.H + 1) + Y.H * (Y.F + 1);
        hipLaunchKernelGGL(k_kan_finish, dim3((units + 255) / 256), dim3(256), 0, s, Y, *params,
                           *grads, b.tab, sums);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kfail(DDRS_ERR_HIP, std::string("kan bwd: ") + hipGetErrorString(e));
    return DDRS_OK;
}

}  // extern "C"

// MI355X (gfx950) kernels for the differentiable Muskingum-Cunge timestep.
//
// Layout. Reaches are renumbered into a level-major schedule: position p
// orders reaches by (level, reach id), where level is the longest upstream
// path. Every upstream reach of p therefore sits at a smaller position, and a
// 64-reach chunk (one wavefront, one workgroup) only depends on chunks at lower
// levels plus, where a level holds fewer than 64 reaches, on earlier sub-levels
// of itself. Per-reach statics and state live in schedule order in HBM.
//
// Synchronisation. A chunk publishes each solved value as a tagged 64-bit word
// (fp32 bits | launch epoch << 32) with one relaxed agent-scope atomic store;
// consumers poll with relaxed agent-scope atomic loads and accept a word only
// when its tag is the current epoch. One single-copy-atomic word carries value
// and readiness together, so no release/acquire fence (and no L2 writeback) is
// needed per hop. Within a chunk, values pass through LDS level by level.
//
// Time skew. A persistent launch owns chunks w, w+W, ... and advances each one
// through the window's timesteps independently, polling without blocking, so
// chunk c can be at step t+5 while a downstream chunk is still at step t: the
// critical path becomes (chunk levels + T) hops instead of (chunk levels x T).
// The grid never exceeds the number of co-resident workgroups, which together
// with non-blocking polling guarantees progress: the least-advanced, lowest
// chunk always has its inputs. A wall-clock watchdog (s_memrealtime) ends every
// wave with an error flag if no chunk of the wave progresses for two seconds.
//
// Arithmetic follows the reference op for op (see mc_physics.h); upstream sums
// are taken in the reference CSR order, so results do not depend on scheduling.

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "mc_physics.h"
#include "ddrs_internal.h"

namespace ddrs {

// ---------------------------------------------------------------------------
// tagged words

__device__ __forceinline__ uint64_t tag_word(float v, uint32_t epoch) {
    return (static_cast<uint64_t>(epoch) << 32) | static_cast<uint64_t>(__float_as_uint(v));
}
__device__ __forceinline__ void publish(uint64_t* slot, float v, uint32_t epoch) {
    __hip_atomic_store(slot, tag_word(v, epoch), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t poll(const uint64_t* slot) {
    return __hip_atomic_load(const_cast<uint64_t*>(slot), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float word_value(uint64_t w) {
    return __uint_as_float(static_cast<uint32_t>(w));
}
__device__ __forceinline__ bool word_ready(uint64_t w, uint32_t epoch) {
    return static_cast<uint32_t>(w >> 32) == epoch;
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// 100 MHz wall clock
__device__ __forceinline__ uint64_t wall_ticks() { return __builtin_amdgcn_s_memrealtime(); }
constexpr uint64_t kWatchdogTicks = 200000000ull;  // 2 s at 100 MHz

// ---------------------------------------------------------------------------
// Persistent driver shared by all solves. Dir = +1 walks steps forward and
// chunks in increasing order, Dir = -1 walks steps backward (step index
// nsteps-1 first) and chunks in decreasing order. For chunk c at window step s:
//   prepare(c, s, slot)  everything that does not depend on other chunks at
//                        step s (coefficients, right-hand side), kept in LDS
//                        slot `slot`;
//   ready(c, s)          polls the cross-chunk inputs of step s and stages them,
//                        with each entry's weight and kind, in LDS;
//   solve(c, s, slot)    the dependent part: the level-by-level solve inside the
//                        chunk, publishing, and the step's epilogue.
// A wave's first kPrepSlots chunks keep a prepared step in LDS, so their
// coefficients are computed while the inputs are still in flight; other
// chunks prepare just in time in the scratch slot kPrepSlots. All three are
// called by the whole wave.
constexpr int kPrepSlots = 2;

template <int Dir, class Prepare, class Ready, class Solve>
__device__ __forceinline__ void persistent_loop(int C, int nsteps, int* tcur, int* pstep,
                                                int* err, Prepare prepare, Ready ready,
                                                Solve solve) {
    const int W = gridDim.x, w = blockIdx.x;
    const int m = (C - w + W - 1) / W;
    for (int k = threadIdx.x; k < m; k += 64) tcur[k] = 0;
    if (threadIdx.x < kPrepSlots) pstep[threadIdx.x] = -1;
    __syncthreads();
    int remaining = m;
    uint64_t last = wall_ticks();
    while (remaining > 0) {
        bool progress = false;
        for (int kk = 0; kk < m; ++kk) {
            const int k = (Dir > 0) ? kk : (m - 1 - kk);
            int done = uni(tcur[k]);
            if (done >= nsteps) continue;
            const int c = w + k * W;
            const int slot = (k < kPrepSlots) ? k : kPrepSlots;
            int burst = 0;
            const bool jit = slot == kPrepSlots;
            while (done < nsteps && burst < 8) {
                const int s = (Dir > 0) ? done : (nsteps - 1 - done);
                // slotted chunks prepare before polling, the scratch slot after
                if (jit && !ready(c, s)) break;
                if (jit || uni(pstep[slot]) != s) {
                    prepare(c, s, slot);
                    __syncthreads();
                    if (!jit && threadIdx.x == 0) pstep[slot] = s;
                    __syncthreads();
                }
                if (!jit && !ready(c, s)) break;
                solve(c, s, slot);
                ++done;
                ++burst;
            }
            if (burst > 0) {
                progress = true;
                __syncthreads();
                if (threadIdx.x == 0) tcur[k] = done;
                __syncthreads();
                if (done >= nsteps) --remaining;
            }
        }
        if (progress) {
            last = wall_ticks();
        } else {
            __builtin_amdgcn_s_sleep(4);
            if (wall_ticks() - last > kWatchdogTicks) {
                if (threadIdx.x == 0) atomicExch(err, 1);
                return;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// forward window

template <int PH>
__global__ __launch_bounds__(64, 5) void k_fwd_window(FwdArgs a) {
    __shared__ float lds_x[64];
    __shared__ float sval[kStageMax], sadj[kStageMax];
    __shared__ short skind[kStageMax];  // >= 0: lane of an intra-chunk upstream; -1: staged
    __shared__ float pb[kPrepSlots + 1][64], pnc1[kPrepSlots + 1][64];
    __shared__ int pe0[kPrepSlots + 1][64], pe1[kPrepSlots + 1][64], porig[kPrepSlots + 1][64];
    __shared__ int psub[kPrepSlots + 1][64];
    __shared__ int pstep[kPrepSlots];
    __shared__ int tcur[kMaxChunksPerWave];
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R;

    // S1..S25 for chunk c at window step s: coefficients and right-hand side.
    auto prepare = [&](int c, int s, int slot) {
        const int p = c * 64 + lane;
        const int base = sc.up_ptr[c * 64];
        if (p < R) {
            const int t = a.t0 + a.s_base + s;
            const int orig = sc.perm[p];
            const float q = a.st_q[p];
            const float it = a.st_i[p];
            const float qp = fmaxr(a.qprime[static_cast<size_t>(t) * R + orig], a.P.q_lb);
            Coef co;
            mc_coef<PH>(a.P, a.st.n[p], a.st.qsp[p], a.st.p[p], q, a.st.len[p], a.st.slope[p],
                        a.st.sqrt_s[p], a.st.xst[p], co);
            pb[slot][lane] = ((co.c2 * it) + (co.c3 * q)) + (co.c4 * qp);        // S25
            pnc1[slot][lane] = -co.c1;
            pe0[slot][lane] = sc.up_ptr[p] - base;
            pe1[slot][lane] = sc.up_ptr[p + 1] - base;
            porig[slot][lane] = orig;
            psub[slot][lane] = sc.sub[p];
        } else {
            psub[slot][lane] = -1;
        }
    };

    auto ready = [&](int c, int s) -> bool {
        const int p0 = c * 64;
        const int p1 = min(p0 + 64, R);
        const int base = sc.up_ptr[p0], end = sc.up_ptr[p1];
        const uint64_t* row = a.hist + static_cast<size_t>(a.s_base + s) * R;
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int u = sc.up_pos[e];
            sadj[e - base] = sc.up_adj[e];
            if ((u >> 6) == c) {
                skind[e - base] = static_cast<short>(u & 63);
            } else {
                skind[e - base] = -1;
                const uint64_t wv = poll(row + u);
                if (!word_ready(wv, a.epoch)) ok = false;
                else sval[e - base] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };

    // S26..S28 and the next step's S24 (N Q_{t+1}) from LDS only.
    auto solve = [&](int c, int s, int slot) {
        const int p = c * 64 + lane;
        const int sg = a.s_base + s;
        const int t = a.t0 + sg;
        uint64_t* row = a.hist + static_cast<size_t>(sg) * R;
        const int my = psub[slot][lane];
        const float b = pb[slot][lane];
        const float nc1 = pnc1[slot][lane];
        const int e0 = pe0[slot][lane], e1 = pe1[slot][lane];
        float x = 0.f, inext = 0.f;
        const int nsub = uni(sc.nsub[c]);
        for (int l = 0; l < nsub; ++l) {
            if (my == l) {
                float sum = 0.0f, in = 0.0f;
                for (int e = e0; e < e1; ++e) {
                    const int kd = skind[e];
                    const float v = (kd >= 0) ? lds_x[kd] : sval[e];
                    const float adj = sadj[e];
                    sum = sum + (nc1 * adj) * v;
                    in = in + fmaxr(v, a.P.q_lb) * adj;
                }
                const float ad = 1.0f + nc1 * 0.0f;
                x = (b - sum) / ad;
                inext = in;
                lds_x[lane] = x;
                publish(row + p, x, a.epoch);
            }
            __syncthreads();
        }
        if (my >= 0) {
            const float qn = fmaxr(x, a.P.q_lb);
            const int orig = porig[slot][lane];
            a.st_q[p] = qn;
            a.st_i[p] = inext;
            if (a.Q_out) a.Q_out[static_cast<size_t>(t + 1) * R + orig] = qn;
            if (a.X_out) a.X_out[orig] = x;
        }
    };

    persistent_loop<+1>(sc.C, a.nsteps, tcur, pstep, a.err, prepare, ready, solve);
}

// ---------------------------------------------------------------------------
// adjoint window

template <int PH>
__global__ __launch_bounds__(64, 4) void k_adj_window(AdjArgs a) {
    __shared__ float ldsP[64], ldsG[64];
    __shared__ float stP[kStageMax], stG[kStageMax], sadj[kStageMax];
    __shared__ short skind[kStageMax];
    __shared__ float pgx[kPrepSlots + 1][64], pnc1[kPrepSlots + 1][64], pc2[kPrepSlots + 1][64];
    __shared__ int pe0[kPrepSlots + 1][64], pe1[kPrepSlots + 1][64], psub[kPrepSlots + 1][64];
    __shared__ int pstep[kPrepSlots];
    __shared__ int tcur[kMaxChunksPerWave];
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R;

    // Everything of reverse step s that does not wait for downstream messages:
    // out[t], N out[t] and q' of the step, the coefficients, and B28.
    auto prepare = [&](int c, int s, int slot) {
        const int p = c * 64 + lane;
        const int base = sc.dn_ptr[c * 64];
        if (p < R) {
            const int sg = a.s_base + s;
            const int t = a.t0 + sg;
            const int orig = sc.perm[p];
            const float xs = word_value(a.hist[static_cast<size_t>(sg) * R + p]);
            float q, it;
            if (sg > 0) {
                const uint64_t* hq = a.hist + static_cast<size_t>(sg - 1) * R;
                q = fmaxr(word_value(hq[p]), a.P.q_lb);
                float acc = 0.0f;
                for (int e = sc.up_ptr[p]; e < sc.up_ptr[p + 1]; ++e)
                    acc = acc + fmaxr(word_value(hq[sc.up_pos[e]]), a.P.q_lb) * sc.up_adj[e];
                it = acc;
            } else {
                q = a.ck_q[p];
                it = a.ck_i[p];
            }
            const float qp_raw = a.qprime[static_cast<size_t>(t) * R + orig];
            Coef co;
            mc_coef<PH>(a.P, a.st.n[p], a.st.qsp[p], a.st.p[p], q, a.st.len[p], a.st.slope[p],
                        a.st.sqrt_s[p], a.st.xst[p], co);
            pgx[slot][lane] = (xs > a.P.q_lb) ? a.st_lam[p] : 0.0f;            // B28
            pnc1[slot][lane] = -co.c1;
            pc2[slot][lane] = co.c2;
            pe0[slot][lane] = sc.dn_ptr[p] - base;
            pe1[slot][lane] = sc.dn_ptr[p + 1] - base;
            psub[slot][lane] = sc.subr[p];
            a.sc_q[p] = q;
            a.sc_i[p] = it;
            a.sc_qp[p] = qp_raw;
        } else {
            psub[slot][lane] = -1;
        }
    };

    auto ready = [&](int c, int s) -> bool {
        const int p0 = c * 64;
        const int p1 = min(p0 + 64, R);
        const int base = sc.dn_ptr[p0], end = sc.dn_ptr[p1];
        const uint64_t* row = a.amsg + static_cast<size_t>(a.s_base + s) * R * 2;
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int d = sc.dn_pos[e];
            sadj[e - base] = sc.dn_adj[e];
            if ((d >> 6) == c) {
                skind[e - base] = static_cast<short>(d & 63);
            } else {
                skind[e - base] = -1;
                const uint64_t w0 = poll(row + 2 * d);
                const uint64_t w1 = poll(row + 2 * d + 1);
                if (!word_ready(w0, a.epoch) || !word_ready(w1, a.epoch)) {
                    ok = false;
                } else {
                    stP[e - base] = word_value(w0);
                    stG[e - base] = word_value(w1);
                }
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };

    auto solve = [&](int c, int s, int slot) {
        const int p = c * 64 + lane;
        const int sg = a.s_base + s;
        const int t = a.t0 + sg;
        uint64_t* mrow = a.amsg + static_cast<size_t>(sg) * R * 2;
        const int my = psub[slot][lane];
        const float gx = pgx[slot][lane];
        const float nc1 = pnc1[slot][lane];
        const float c2 = pc2[slot][lane];
        const int e0 = pe0[slot][lane], e1 = pe1[slot][lane];
        float gb = 0.f, gq24 = 0.f;
        const int nsub = uni(sc.nsubr[c]);
        for (int l = 0; l < nsub; ++l) {
            if (my == l) {                                               // B27 transposed solve
                float sumP = 0.0f, sumG = 0.0f;
                for (int e = e0; e < e1; ++e) {
                    const int kd = skind[e];
                    const float Pd = (kd >= 0) ? ldsP[kd] : stP[e];
                    const float Gd = (kd >= 0) ? ldsG[kd] : stG[e];
                    const float adj = sadj[e];
                    sumP = sumP + adj * Pd;
                    sumG = sumG + Gd * adj;                              // B24 N^T gi
                }
                const float ad = 1.0f + nc1 * 0.0f;
                gb = (gx - sumP) / ad;
                gq24 = sumG;
                const float Pm = nc1 * gb;
                const float Gm = c2 * gb;
                ldsP[lane] = Pm;
                ldsG[lane] = Gm;
                publish(mrow + 2 * p, Pm, a.epoch);
                publish(mrow + 2 * p + 1, Gm, a.epoch);
            }
            __syncthreads();
        }
        if (my >= 0) {
            const uint64_t* hx = a.hist + static_cast<size_t>(sg) * R;
            // B26: gc1 = -sum_k adj_k * ((-gb) * x[col k]) in CSR order
            float acc = 0.0f;
            for (int e = sc.up_ptr[p]; e < sc.up_ptr[p + 1]; ++e) {
                const float gA = (-gb) * word_value(hx[sc.up_pos[e]]);
                acc = acc + gA * sc.up_adj[e];
            }
            const float gc1 = -acc;
            const float q = a.sc_q[p], it = a.sc_i[p], qp_raw = a.sc_qp[p];
            const float qp = fmaxr(qp_raw, a.P.q_lb);
            Coef co;
            mc_coef<PH>(a.P, a.st.n[p], a.st.qsp[p], a.st.p[p], q, a.st.len[p], a.st.slope[p],
                        a.st.sqrt_s[p], a.st.xst[p], co);
            const LocalGrads g = mc_local_bwd<PH>(a.P, co, a.st.n[p], a.st.p[p], q, qp,
                                                  a.st.len[p], a.st.slope[p], a.st.sqrt_s[p],
                                                  it, gb, gc1, gq24);
            a.acc_n[p] = a.acc_n[p] + g.gn;
            a.acc_qsp[p] = a.acc_qsp[p] + g.gqsp;
            a.acc_p[p] = a.acc_p[p] + g.gp;
            const int orig = sc.perm[p];
            const float direct = a.gQ ? a.gQ[static_cast<size_t>(t) * R + orig] : 0.0f;
            a.st_lam[p] = direct + g.gq;
            if (a.gqp_out)
                a.gqp_out[static_cast<size_t>(t) * R + orig] =
                    (!a.mask_qp || qp_raw > a.P.q_lb) ? g.gqp : 0.0f;
        }
    };

    persistent_loop<-1>(sc.C, a.nsteps, tcur, pstep, a.err, prepare, ready, solve);
}

// ---------------------------------------------------------------------------
// general triangular solves (one step)

__global__ __launch_bounds__(64) void k_tri_fwd(TriArgs a) {
    __shared__ float lds_x[64];
    __shared__ float stage[kStageMax];
    __shared__ int tcur[kMaxChunksPerWave];
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R;
    auto ready = [&](int c, int) -> bool {
        const int p0 = c * 64, p1 = min(p0 + 64, R);
        const int base = sc.up_ptr[p0], end = sc.up_ptr[p1];
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int u = sc.up_pos[e];
            if ((u >> 6) != c) {
                const uint64_t wv = poll(a.slot + u);
                if (!word_ready(wv, a.epoch)) ok = false;
                else stage[e - base] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };
    auto process = [&](int c, int) {
        const int p = c * 64 + lane;
        const bool valid = p < R;
        const int base = sc.up_ptr[c * 64];
        int my = valid ? sc.sub[p] : -1;
        float x = 0.f;
        const int nsub = uni(sc.nsub[c]);
        for (int l = 0; l < nsub; ++l) {
            if (my == l) {
                float sum = 0.0f;
                for (int e = sc.up_ptr[p]; e < sc.up_ptr[p + 1]; ++e) {
                    const int u = sc.up_pos[e];
                    const float v = ((u >> 6) == c) ? lds_x[u & 63] : stage[e - base];
                    const float av = a.a_values ? a.a_values[sc.up_k[e]] : -sc.up_adj[e];
                    sum = sum + av * v;
                }
                const float diag = a.a_values ? a.a_values[sc.diag_k[p]] : 1.0f;
                x = (a.b[sc.perm[p]] - sum) / diag;
                lds_x[lane] = x;
                publish(a.slot + p, x, a.epoch);
            }
            __syncthreads();
        }
        if (valid) a.x[sc.perm[p]] = a.clamp ? fmaxr(x, a.lb) : x;
    };
    __shared__ int pstep[kPrepSlots];
    auto prepare = [&](int, int, int) {};
    auto solve = [&](int c, int s, int) { process(c, s); };
    persistent_loop<+1>(sc.C, 1, tcur, pstep, a.err, prepare, ready, solve);
}

__global__ __launch_bounds__(64) void k_tri_adj(TriArgs a) {
    __shared__ float lds_g[64];
    __shared__ float stage[kStageMax];
    __shared__ int tcur[kMaxChunksPerWave];
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R;
    auto ready = [&](int c, int) -> bool {
        const int p0 = c * 64, p1 = min(p0 + 64, R);
        const int base = sc.dn_ptr[p0], end = sc.dn_ptr[p1];
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int d = sc.dn_pos[e];
            if ((d >> 6) != c) {
                const uint64_t wv = poll(a.slot + d);
                if (!word_ready(wv, a.epoch)) ok = false;
                else stage[e - base] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };
    auto process = [&](int c, int) {
        const int p = c * 64 + lane;
        const bool valid = p < R;
        const int base = sc.dn_ptr[c * 64];
        int my = valid ? sc.subr[p] : -1;
        float g = 0.f;
        const int nsub = uni(sc.nsubr[c]);
        for (int l = 0; l < nsub; ++l) {
            if (my == l) {
                float sum = 0.0f;
                for (int e = sc.dn_ptr[p]; e < sc.dn_ptr[p + 1]; ++e) {
                    const int d = sc.dn_pos[e];
                    const float v = ((d >> 6) == c) ? lds_g[d & 63] : stage[e - base];
                    const float av = a.a_values ? a.a_values[sc.dn_k[e]] : -sc.dn_adj[e];
                    sum = sum + av * v;
                }
                const float diag = a.a_values ? a.a_values[sc.diag_k[p]] : 1.0f;
                const int o = sc.perm[p];
                const float gin = a.mask_x ? ((a.mask_x[o] > a.lb) ? a.b[o] : 0.0f) : a.b[o];
                g = (gin - sum) / diag;
                lds_g[lane] = g;
                publish(a.slot + p, g, a.epoch);
            }
            __syncthreads();
        }
        if (valid) a.x[sc.perm[p]] = g;
    };
    __shared__ int pstep[kPrepSlots];
    auto prepare = [&](int, int, int) {};
    auto solve = [&](int c, int s, int) { process(c, s); };
    persistent_loop<-1>(sc.C, 1, tcur, pstep, a.err, prepare, ready, solve);
}

// grad_a[k] = (-gb[row k]) * x[col k] over the CSR entries (dispatch.rs:120-136)
__global__ void k_grad_a(int nnz, const int* row_for_nnz, const int* col, const float* gb,
                         const float* x, float* grad_a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) grad_a[k] = (-gb[row_for_nnz[k]]) * x[col[k]];
}

// ---------------------------------------------------------------------------
// gathers between reference order and schedule order

__global__ void k_bind_statics(int R, const int* perm, const float* n, const float* qsp,
                               const float* psp, const float* len, const float* slope,
                               const float* xst, float slope_lb, Statics st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    const int o = perm[p];
    st.n[p] = n[o];
    st.qsp[p] = qsp[o];
    st.p[p] = psp[o];
    st.len[p] = len[o];
    const float s = fmaxr(slope[o], slope_lb);  // setup_inputs clamp_min (mmc.rs:259-264)
    st.slope[p] = s;
    st.sqrt_s[p] = sqrtf(s);
    st.xst[p] = xst ? xst[o] : 0.3f;
}

// st_q[p] = (clamp ? max(q[o], lb) : q[o]); optionally Q_out row 0.
__global__ void k_state_q(int R, const int* perm, const float* q, int clamp, float lb,
                          float* st_q, float* Q_row0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    const int o = perm[p];
    const float v = clamp ? fmaxr(q[o], lb) : q[o];
    st_q[p] = v;
    if (Q_row0) Q_row0[o] = v;
}

// st_i[p] = sum_k st_q[up k] * adj_k in CSR order (S24 of the first step).
__global__ void k_state_i(Sched sc, const float* st_q, float* st_i) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.R) return;
    float acc = 0.0f;
    for (int e = sc.up_ptr[p]; e < sc.up_ptr[p + 1]; ++e) acc = acc + st_q[sc.up_pos[e]] * sc.up_adj[e];
    st_i[p] = acc;
}

// x history row 0 from a caller vector (single-step backward).
__global__ void k_hist_from_x(int R, const int* perm, const float* x, uint64_t* hist) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    hist[p] = tag_word(x[perm[p]], 0u);
}

__global__ void k_lam_init(int R, const int* perm, const float* g, float* lam, float* a0,
                           float* a1, float* a2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    lam[p] = g ? g[perm[p]] : 0.0f;
    a0[p] = 0.0f;
    a1[p] = 0.0f;
    a2[p] = 0.0f;
}

__global__ void k_unpermute(int R, const int* perm, const float* src, float* dst) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    dst[perm[p]] = src[p];
}

// dL/dQ0 through out[0] = max(Q0, lb) (mmc.rs:536-541)
__global__ void k_grad_q0(int R, const int* perm, const float* lam, const float* q0, float lb,
                          float* g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R) return;
    const int o = perm[p];
    g[o] = (q0[o] > lb) ? lam[p] : 0.0f;
}

__global__ void k_zero_f(size_t n, float* x) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// launch helpers

template <int PH>
static void launch_fwd_t(const FwdArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_fwd_window<PH>, dim3(grid), dim3(64), 0, s, a);
}
template <int PH>
static void launch_adj_t(const AdjArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_window<PH>, dim3(grid), dim3(64), 0, s, a);
}

static int phys_variant(const ddrs_params* p) {
    if (p->ddr_match) return PH_DDR;
    if (p->enforce_positivity) return PH_POS;
    return PH_DEFAULT;
}

void launch_fwd(const FwdArgs& a, const ddrs_params* p, int grid, hipStream_t s) {
    switch (phys_variant(p)) {
        case PH_DDR: launch_fwd_t<PH_DDR>(a, grid, s); break;
        case PH_POS: launch_fwd_t<PH_POS>(a, grid, s); break;
        default: launch_fwd_t<PH_DEFAULT>(a, grid, s); break;
    }
}
void launch_adj(const AdjArgs& a, const ddrs_params* p, int grid, hipStream_t s) {
    switch (phys_variant(p)) {
        case PH_DDR: launch_adj_t<PH_DDR>(a, grid, s); break;
        case PH_POS: launch_adj_t<PH_POS>(a, grid, s); break;
        default: launch_adj_t<PH_DEFAULT>(a, grid, s); break;
    }
}
void launch_tri_fwd(const TriArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_fwd, dim3(grid), dim3(64), 0, s, a);
}
void launch_tri_adj(const TriArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_adj, dim3(grid), dim3(64), 0, s, a);
}

static inline dim3 g1(size_t n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

void launch_grad_a(int nnz, const int* row, const int* col, const float* gb, const float* x,
                   float* ga, hipStream_t s) {
    hipLaunchKernelGGL(k_grad_a, g1(nnz), dim3(256), 0, s, nnz, row, col, gb, x, ga);
}
void launch_bind_statics(int R, const int* perm, const float* n, const float* qsp,
                         const float* psp, const float* len, const float* slope,
                         const float* xst, float slope_lb, const Statics& st, hipStream_t s) {
    hipLaunchKernelGGL(k_bind_statics, g1(R), dim3(256), 0, s, R, perm, n, qsp, psp, len, slope,
                       xst, slope_lb, st);
}
void launch_state(const Sched& sc, const float* q, int clamp, float lb, float* st_q,
                  float* st_i, float* Q_row0, hipStream_t s) {
    hipLaunchKernelGGL(k_state_q, g1(sc.R), dim3(256), 0, s, sc.R, sc.perm, q, clamp, lb, st_q,
                       Q_row0);
    hipLaunchKernelGGL(k_state_i, g1(sc.R), dim3(256), 0, s, sc, st_q, st_i);
}
void launch_hist_from_x(int R, const int* perm, const float* x, uint64_t* hist, hipStream_t s) {
    hipLaunchKernelGGL(k_hist_from_x, g1(R), dim3(256), 0, s, R, perm, x, hist);
}
void launch_lam_init(int R, const int* perm, const float* g, float* lam, float* a0, float* a1,
                     float* a2, hipStream_t s) {
    hipLaunchKernelGGL(k_lam_init, g1(R), dim3(256), 0, s, R, perm, g, lam, a0, a1, a2);
}
void launch_unpermute(int R, const int* perm, const float* src, float* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_unpermute, g1(R), dim3(256), 0, s, R, perm, src, dst);
}
void launch_grad_q0(int R, const int* perm, const float* lam, const float* q0, float lb,
                    float* g, hipStream_t s) {
    hipLaunchKernelGGL(k_grad_q0, g1(R), dim3(256), 0, s, R, perm, lam, q0, lb, g);
}
void launch_zero(size_t n, float* x, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_zero_f, g1(n), dim3(256), 0, s, n, x);
}

int occupancy_blocks(int which) {
    int nb = 0;
    hipError_t e = hipSuccess;
    switch (which) {
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fwd_window<PH_DEFAULT>, 64, 0); break;
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fwd_window<PH_DDR>, 64, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fwd_window<PH_POS>, 64, 0); break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj_window<PH_DEFAULT>, 64, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj_window<PH_DDR>, 64, 0); break;
        case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj_window<PH_POS>, 64, 0); break;
        case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tri_fwd, 64, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tri_adj, 64, 0); break;
    }
    return (e == hipSuccess) ? nb : 0;
}

}  // namespace ddrs

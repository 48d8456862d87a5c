// MI355X (gfx950) kernels for the differentiable Muskingum-Cunge timestep.
//
// Layout. Reaches are renumbered into a level-major schedule: position p
// orders reaches by (level, reach id), where level is the longest upstream
// path. Every upstream reach of p therefore sits at a smaller position. The
// schedule is cut into chunks of 64 * RPL consecutive positions (RPL reaches
// per lane, one wavefront per workgroup); a chunk depends only on chunks with
// smaller indices plus, where a level holds fewer reaches than the chunk, on
// earlier sub-levels of itself. Per-reach statics and state live in schedule
// order in HBM.
//
// Synchronisation. A chunk publishes each solved value as a tagged 64-bit word
// (fp32 bits | launch epoch << 32) with one relaxed agent-scope atomic store;
// consumers poll with relaxed agent-scope atomic loads and accept a word only
// when its tag is the current epoch. One single-copy-atomic word carries value
// and readiness together, so no release/acquire fence (and no L2 writeback) is
// needed per hop. Within a chunk, values pass through LDS level by level.
//
// Time skew. A persistent launch has one workgroup per chunk, and each
// workgroup advances its chunk through all of the launch's timesteps, so chunk
// c can be at step t+5 while a downstream chunk is still at step t: the
// critical path becomes (chunk levels + T) hops instead of (chunk levels x T).
// Workgroups take their chunk from an atomic counter when they start, in
// dependency order (upstream first for the forward, downstream first for the
// adjoint), so every chunk a running workgroup waits on belongs to a
// workgroup that has already started: progress does not depend on how many
// workgroups the GPU holds at once. A wall-clock watchdog (s_memrealtime) ends
// a workgroup with an error flag if its inputs do not arrive for two seconds.
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
// One chunk per workgroup, for all steps of the launch.

// Chunk of this workgroup: the claim counter only grows (claims of earlier
// launches are subtracted through `base`), so it is never reset. Dir = +1
// hands out chunks 0, 1, ..., Dir = -1 hands out C-1, C-2, ...
template <int Dir>
__device__ __forceinline__ int claim_chunk(const Claim& cl, int C, int* lds) {
    if (threadIdx.x == 0) {
        const unsigned long long k = atomicAdd(cl.ctr, 1ull) - cl.base;
        *lds = static_cast<int>(k);
    }
    __syncthreads();
    const int k = uni(*lds);
    return (Dir > 0) ? k : (C - 1 - k);
}

// For the launch's steps in time order (Dir = +1) or reverse (Dir = -1):
//   prepare(s, first)  everything of step s that does not depend on other
//                      chunks (coefficients, right-hand side, into LDS);
//                      `first` asks for the chunk's step-invariant data too;
//   ready(s)           polls the cross-chunk inputs of step s into LDS;
//   solve(s, last)     the level-by-level solve inside the chunk, publishing,
//                      and the step's epilogue; `last` is the launch's final step.
// All three are called by the whole wave. Returns early (error flag set) when
// the inputs of a step do not arrive within the watchdog time or another
// workgroup has raised the flag.
template <int Dir, class Prepare, class Ready, class Solve>
__device__ __forceinline__ void chunk_loop(int nsteps, int* err, Prepare prepare, Ready ready,
                                           Solve solve) {
    for (int done = 0; done < nsteps; ++done) {
        const int s = (Dir > 0) ? done : (nsteps - 1 - done);
        prepare(s, done == 0);
        __syncthreads();
        if (!ready(s)) {
            const uint64_t t0 = wall_ticks();
            do {
                __builtin_amdgcn_s_sleep(2);
                const bool give_up =
                    (wall_ticks() - t0 > kWatchdogTicks) ||
                    __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                if (give_up) {
                    if (threadIdx.x == 0) atomicExch(err, 1);
                    return;
                }
            } while (!ready(s));
        }
        solve(s, done == nsteps - 1);
    }
}

template <int RPL>
struct Width {
    static constexpr int CW = 64 * RPL;                  // positions per chunk
    static constexpr int SHIFT = (RPL == 1) ? 6 : 7;
    static constexpr int kStageUp = kStageUpBase * RPL;  // staged upstream links
    static constexpr int kStageDn = kStageDnBase * RPL;  // staged downstream links
};

// Stage the links [base, end) of chunk c into LDS: link position, weight, and
// kind (>= 0: chunk-local index of a link inside the chunk, -1: polled from
// another chunk).
template <int RPL>
__device__ __forceinline__ void stage_links(int c, int base, int end, const int* pos,
                                            const float* w, int* spos, float* sadj,
                                            short* skind) {
    for (int e = base + threadIdx.x; e < end; e += 64) {
        const int u = pos[e];
        spos[e - base] = u;
        sadj[e - base] = w[e];
        skind[e - base] = ((u >> Width<RPL>::SHIFT) == c)
                              ? static_cast<short>(u & (Width<RPL>::CW - 1))
                              : static_cast<short>(-1);
    }
}

// ---------------------------------------------------------------------------
// forward window

template <int PH, int RPL>
__global__ __launch_bounds__(64) void k_fwd_window(FwdArgs a) {
    using Wd = Width<RPL>;
    constexpr int CW = Wd::CW;
    __shared__ float lds_x[CW];
    __shared__ float sval[Wd::kStageUp], sadj[Wd::kStageUp];
    __shared__ int spos[Wd::kStageUp];
    __shared__ short skind[Wd::kStageUp];
    __shared__ float pb[CW], pnc1[CW], sq[CW], si[CW];
    __shared__ short pe0[CW], pe1[CW], psub[CW];
    __shared__ int porig[CW];
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R, P = sc.P;
    const int c = claim_chunk<+1>(a.claim, a.ch.C, &claimed);
    const int p0 = c * CW;
    const int base = sc.up_ptr[p0];
    const int nlinks = sc.up_ptr[p0 + CW] - base;
    const int nsub = a.ch.nsub[c];

    // S1..S25 for window step s: coefficients and right-hand side.
    auto prepare = [&](int s, bool first) {
        if (first) stage_links<RPL>(c, base, base + nlinks, sc.up_pos, sc.up_adj, spos, sadj, skind);
        const int t = a.t0 + a.s_base + s;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int p = p0 + j;
            if (first) porig[j] = sc.perm[p];
            if (porig[j] >= 0) {
                if (first) {
                    pe0[j] = static_cast<short>(sc.up_ptr[p] - base);
                    pe1[j] = static_cast<short>(sc.up_ptr[p + 1] - base);
                    psub[j] = a.ch.sub[p];
                    sq[j] = a.st_q[p];
                    si[j] = a.st_i[p];
                }
                const float q = sq[j];
                const float it = si[j];
                const float qp = fmaxr(a.qprime[static_cast<size_t>(t) * R + porig[j]], a.P.q_lb);
                const Reach rc = load_reach(a.st, p);
                Coef co;
                mc_coef<PH>(a.P, rc, q, co);
                pb[j] = ((co.c2 * it) + (co.c3 * q)) + (co.c4 * qp);             // S25
                pnc1[j] = -co.c1;
            } else if (first) {
                psub[j] = -1;
            }
        }
    };

    auto ready = [&](int s) -> bool {
        const uint64_t* row = a.hist + static_cast<size_t>(a.s_base + s) * P;
        bool ok = true;
        for (int e = lane; e < nlinks; e += 64) {
            if (skind[e] < 0) {
                const uint64_t wv = poll(row + spos[e]);
                if (!word_ready(wv, a.epoch)) ok = false;
                else sval[e] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };

    // S26..S28 and the next step's S24 (N Q_{t+1}) from LDS only.
    auto solve = [&](int s, bool last) {
        const int sg = a.s_base + s;
        const int t = a.t0 + sg;
        uint64_t* row = a.hist + static_cast<size_t>(sg) * P;
        int my[RPL];
        float x[RPL], inext[RPL];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            my[r] = psub[r * 64 + lane];
            x[r] = 0.0f;
            inext[r] = 0.0f;
        }
        for (int l = 0; l < nsub; ++l) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (my[r] == l) {
                    const int j = r * 64 + lane;
                    const float b = pb[j];
                    const float nc1 = pnc1[j];
                    const int e1 = pe1[j];
                    float sum = 0.0f, in = 0.0f;
                    for (int e = pe0[j]; e < e1; ++e) {
                        const int kd = skind[e];
                        const float v = (kd >= 0) ? lds_x[kd] : sval[e];
                        const float adj = sadj[e];
                        sum = sum + (nc1 * adj) * v;
                        in = in + fmaxr(v, a.P.q_lb) * adj;
                    }
                    const float ad = 1.0f + nc1 * 0.0f;
                    x[r] = (b - sum) / ad;
                    inext[r] = in;
                    lds_x[j] = x[r];
                    publish(row + p0 + j, x[r], a.epoch);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            if (my[r] >= 0) {
                const int j = r * 64 + lane;
                const float qn = fmaxr(x[r], a.P.q_lb);
                sq[j] = qn;
                si[j] = inext[r];
                if (last) {
                    a.st_q[p0 + j] = qn;
                    a.st_i[p0 + j] = inext[r];
                }
                if (a.Q_out) a.Q_out[static_cast<size_t>(t + 1) * R + porig[j]] = qn;
                if (a.X_out) a.X_out[porig[j]] = x[r];
            }
        }
    };

    chunk_loop<+1>(a.nsteps, a.err, prepare, ready, solve);
}

// ---------------------------------------------------------------------------
// adjoint window, phase 1: linearisation of every reach-step
//
// One step's backward (B28..B1) is linear in the transposed-solve value gb once
// the forward state is fixed. This kernel evaluates, for every reach and step
// of the segment at once, the coefficients (S1..S23 from the forward history)
// and the backward chain for gb = 1; the reverse-time sweep then only solves
// for gb, and the parameter gradients are gb times these records.

template <int PH>
__global__ __launch_bounds__(256) void k_adj_lin(LinArgs a) {
    const Sched& sc = a.sc;
    const int R = sc.R, P = sc.P;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int orig = sc.perm[p];
    if (orig < 0) return;
    const int sg = a.s_base + static_cast<int>(blockIdx.y);
    const int t = a.t0 + sg;
    const float lb = a.P.q_lb;
    const int e0 = sc.up_ptr[p], e1 = sc.up_ptr[p + 1];
    const uint64_t* hx = a.hist + static_cast<size_t>(sg) * P;
    float q, it;
    if (sg > 0) {
        const uint64_t* hq = hx - P;
        q = fmaxr(word_value(hq[p]), lb);
        float acc = 0.0f;
        for (int e = e0; e < e1; ++e)
            acc = acc + fmaxr(word_value(hq[sc.up_pos[e]]), lb) * sc.up_adj[e];
        it = acc;
    } else {
        q = a.ck_q[p];
        it = a.ck_i[p];
    }
    // B26 for gb = 1: gc1 = -sum_k adj_k * ((-1) * x[col k]) = sum_k x[col k] * adj_k
    float sx = 0.0f;
    for (int e = e0; e < e1; ++e) sx = sx + word_value(hx[sc.up_pos[e]]) * sc.up_adj[e];
    const float qp_raw = a.qprime[static_cast<size_t>(t) * R + orig];
    const float qp = fmaxr(qp_raw, lb);
    const Reach rc = load_reach(a.st, p);
    Coef co;
    mc_coef<PH>(a.P, rc, q, co);
    const LocalGrads g = mc_local_bwd<PH>(a.P, co, rc, q, qp, it, 1.0f, sx, 0.0f);
    const size_t o = static_cast<size_t>(sg) * P + p;
    const float dq = a.gQ ? a.gQ[static_cast<size_t>(t) * R + orig] : 0.0f;
    a.rec[o] = make_float4(-co.c1, co.c2, g.gq, dq);
    a.recm[o] = (!a.mask_lam || q > lb) ? 1 : 0;
    a.recj[o] = make_float4(g.gn, g.gqsp, g.gp, (!a.mask_qp || qp_raw > lb) ? g.gqp : 0.0f);
}

// ---------------------------------------------------------------------------
// adjoint window, phase 2: reverse-time transposed solves (B28, B27, B24)
//
// Reverse step s of reach p: gx = masked dL/d out[t+1]; gb = gx - sum over
// downstream d of adj * P_d; P_p = -c1 gb and G_p = c2 gb go to the upstream
// reaches; dL/d out[t] = direct + (gb * dgq/dgb + sum_d G_d * adj), masked by
// out[t] > lb when it is passed through the clamp of the previous step.

template <int RPL>
__global__ __launch_bounds__(64) void k_adj_sweep(SweepArgs a) {
    using Wd = Width<RPL>;
    constexpr int CW = Wd::CW;
    __shared__ float ldsP[CW], ldsG[CW];
    __shared__ float stP[Wd::kStageDn], stG[Wd::kStageDn], sadj[Wd::kStageDn];
    __shared__ int spos[Wd::kStageDn];
    __shared__ short skind[Wd::kStageDn];
    __shared__ float4 prec[CW];
    __shared__ float slam[CW];
    __shared__ unsigned char pm[CW], pvalid[CW];
    __shared__ short pe0[CW], pe1[CW], psub[CW];
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int P = sc.P;
    const int c = claim_chunk<-1>(a.claim, a.ch.C, &claimed);
    const int p0 = c * CW;
    const int base = sc.dn_ptr[p0];
    const int nlinks = sc.dn_ptr[p0 + CW] - base;
    const int nsub = a.ch.nsubr[c];

    auto prepare = [&](int s, bool first) {
        if (first) stage_links<RPL>(c, base, base + nlinks, sc.dn_pos, sc.dn_adj, spos, sadj, skind);
        const size_t row = static_cast<size_t>(a.s_base + s) * P;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int p = p0 + j;
            if (first) pvalid[j] = sc.perm[p] >= 0;
            if (pvalid[j]) {
                if (first) {
                    pe0[j] = static_cast<short>(sc.dn_ptr[p] - base);
                    pe1[j] = static_cast<short>(sc.dn_ptr[p + 1] - base);
                    psub[j] = a.ch.subr[p];
                    slam[j] = a.st_lam[p];
                }
                prec[j] = a.rec[row + p];
                pm[j] = a.recm[row + p];
            } else if (first) {
                psub[j] = -1;
            }
        }
    };

    auto ready = [&](int s) -> bool {
        const uint64_t* row = a.amsg + static_cast<size_t>(a.s_base + s) * P * 2;
        bool ok = true;
        for (int e = lane; e < nlinks; e += 64) {
            if (skind[e] < 0) {
                const int d = spos[e];
                const uint64_t w0 = poll(row + 2 * d);
                const uint64_t w1 = poll(row + 2 * d + 1);
                if (!word_ready(w0, a.epoch) || !word_ready(w1, a.epoch)) {
                    ok = false;
                } else {
                    stP[e] = word_value(w0);
                    stG[e] = word_value(w1);
                }
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };

    auto solve = [&](int s, bool last) {
        const int sg = a.s_base + s;
        uint64_t* mrow = a.amsg + static_cast<size_t>(sg) * P * 2;
        int my[RPL];
        float gbv[RPL], lamv[RPL];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            my[r] = psub[r * 64 + lane];
            gbv[r] = 0.0f;
            lamv[r] = 0.0f;
        }
        for (int l = 0; l < nsub; ++l) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (my[r] == l) {
                    const int j = r * 64 + lane;
                    const float4 rc = prec[j];  // {-c1, c2, dgq/dgb, direct}
                    const int e1 = pe1[j];
                    float sumP = 0.0f, sumG = 0.0f;
                    for (int e = pe0[j]; e < e1; ++e) {
                        const int kd = skind[e];
                        const float Pd = (kd >= 0) ? ldsP[kd] : stP[e];
                        const float Gd = (kd >= 0) ? ldsG[kd] : stG[e];
                        const float adj = sadj[e];
                        sumP = sumP + adj * Pd;                          // B27
                        sumG = sumG + Gd * adj;                          // B24 N^T gi
                    }
                    const float ad = 1.0f + rc.x * 0.0f;
                    const float gb = (slam[j] - sumP) / ad;
                    const float Pm = rc.x * gb;
                    const float Gm = rc.y * gb;
                    ldsP[j] = Pm;
                    ldsG[j] = Gm;
                    publish(mrow + 2 * (p0 + j), Pm, a.epoch);
                    publish(mrow + 2 * (p0 + j) + 1, Gm, a.epoch);
                    gbv[r] = gb;
                    lamv[r] = pm[j] ? rc.w + (gb * rc.z + sumG) : 0.0f;
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            if (my[r] >= 0) {
                const int j = r * 64 + lane;
                slam[j] = lamv[r];
                if (last) a.st_lam[p0 + j] = lamv[r];
                a.gbuf[static_cast<size_t>(sg) * P + p0 + j] = gbv[r];
            }
        }
    };

    chunk_loop<-1>(a.nsteps, a.err, prepare, ready, solve);
}

// ---------------------------------------------------------------------------
// adjoint window, phase 3: gradients of n, q_spatial, p_spatial (accumulated
// over steps in reverse time order) and of q'.

__global__ __launch_bounds__(256) void k_adj_fin(FinArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const int orig = a.perm[p];
    if (orig < 0) return;
    float an = a.acc_n[p], aq = a.acc_qsp[p], ap = a.acc_p[p];
    for (int s = a.nsteps - 1; s >= 0; --s) {
        const size_t o = static_cast<size_t>(a.s_base + s) * a.P + p;
        const float gb = a.gbuf[o];
        const float4 j = a.recj[o];
        an = an + gb * j.x;
        aq = aq + gb * j.y;
        ap = ap + gb * j.z;
        if (a.gqp_out)
            a.gqp_out[static_cast<size_t>(a.t0 + a.s_base + s) * a.R + orig] = gb * j.w;
    }
    a.acc_n[p] = an;
    a.acc_qsp[p] = aq;
    a.acc_p[p] = ap;
}

// ---------------------------------------------------------------------------
// general triangular solves (one step, 64-wide chunks)

__global__ __launch_bounds__(64) void k_tri_fwd(TriArgs a) {
    __shared__ float lds_x[64];
    __shared__ float stage[kStageUpBase];
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int c = claim_chunk<+1>(a.claim, a.ch.C, &claimed);
    const int base = sc.up_ptr[c * 64];
    const int end = sc.up_ptr[c * 64 + 64];
    auto prepare = [&](int, bool) {};
    auto ready = [&](int) -> bool {
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
    auto solve = [&](int, bool) {
        const int p = c * 64 + lane;
        const bool valid = sc.perm[p] >= 0;
        int my = valid ? a.ch.sub[p] : -1;
        float x = 0.f;
        const int nsub = a.ch.nsub[c];
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
    chunk_loop<+1>(1, a.err, prepare, ready, solve);
}

__global__ __launch_bounds__(64) void k_tri_adj(TriArgs a) {
    __shared__ float lds_g[64];
    __shared__ float stage[kStageDnBase];
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int c = claim_chunk<-1>(a.claim, a.ch.C, &claimed);
    const int base = sc.dn_ptr[c * 64];
    const int end = sc.dn_ptr[c * 64 + 64];
    auto prepare = [&](int, bool) {};
    auto ready = [&](int) -> bool {
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
    auto solve = [&](int, bool) {
        const int p = c * 64 + lane;
        const bool valid = sc.perm[p] >= 0;
        int my = valid ? a.ch.subr[p] : -1;
        float g = 0.f;
        const int nsub = a.ch.nsubr[c];
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
    chunk_loop<-1>(1, a.err, prepare, ready, solve);
}

// grad_a[k] = (-gb[row k]) * x[col k] over the CSR entries (dispatch.rs:120-136)
__global__ void k_grad_a(int nnz, const int* row_for_nnz, const int* col, const float* gb,
                         const float* x, float* grad_a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) grad_a[k] = (-gb[row_for_nnz[k]]) * x[col[k]];
}

// ---------------------------------------------------------------------------
// gathers between reference order and schedule order (over positions p < P;
// empty positions hold zeros)

__global__ void k_bind_statics(Sched sc, const float* n, const float* qsp, const float* psp,
                               const float* len, const float* slope, const float* xst,
                               float slope_lb, float4* st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    if (o < 0) return;
    const float sl = fmaxr(slope[o], slope_lb);  // setup_inputs clamp_min (mmc.rs:259-264)
    make_reach(n[o], qsp[o], psp[o], len[o], sl, xst ? xst[o] : 0.3f, st + 3 * static_cast<size_t>(p));
}

// st_q[p] = (clamp ? max(q[o], lb) : q[o]); optionally Q_out row 0.
__global__ void k_state_q(Sched sc, const float* q, int clamp, float lb, float* st_q,
                          float* Q_row0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    if (o < 0) {
        st_q[p] = 0.0f;
        return;
    }
    const float v = clamp ? fmaxr(q[o], lb) : q[o];
    st_q[p] = v;
    if (Q_row0) Q_row0[o] = v;
}

// st_i[p] = sum_k st_q[up k] * adj_k in CSR order (S24 of the first step).
__global__ void k_state_i(Sched sc, const float* st_q, float* st_i) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    float acc = 0.0f;
    for (int e = sc.up_ptr[p]; e < sc.up_ptr[p + 1]; ++e) acc = acc + st_q[sc.up_pos[e]] * sc.up_adj[e];
    st_i[p] = acc;
}

// x history row 0 from a caller vector (single-step backward).
__global__ void k_hist_from_x(Sched sc, const float* x, uint64_t* hist) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    hist[p] = tag_word(o >= 0 ? x[o] : 0.0f, 0u);
}

// Gradient entering the last step's solve: dL/d out[T-1] through max(x, lb)
// (B28), with x the last row of the forward history.
__global__ void k_lam_init(Sched sc, const float* g, const uint64_t* hist_row, float lb,
                           float* lam, float* a0, float* a1, float* a2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    const float gv = (g && o >= 0) ? g[o] : 0.0f;
    lam[p] = (!hist_row || word_value(hist_row[p]) > lb) ? gv : 0.0f;
    a0[p] = 0.0f;
    a1[p] = 0.0f;
    a2[p] = 0.0f;
}

__global__ void k_unpermute(Sched sc, const float* src, float* dst) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    if (o >= 0) dst[o] = src[p];
}

// dL/dQ0 through out[0] = max(Q0, lb) (mmc.rs:536-541)
__global__ void k_grad_q0(Sched sc, const float* lam, const float* q0, float lb, float* g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sc.P) return;
    const int o = sc.perm[p];
    if (o >= 0) g[o] = (q0[o] > lb) ? lam[p] : 0.0f;
}

__global__ void k_zero_f(size_t n, float* x) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// launch helpers

template <int PH, int RPL>
static void launch_fwd_t(const FwdArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_fwd_window<PH, RPL>), dim3(grid), dim3(64), 0, s, a);
}
template <int RPL>
static void launch_fwd_r(const FwdArgs& a, int variant, int grid, hipStream_t s) {
    switch (variant) {
        case PH_DDR: launch_fwd_t<PH_DDR, RPL>(a, grid, s); break;
        case PH_POS: launch_fwd_t<PH_POS, RPL>(a, grid, s); break;
        default: launch_fwd_t<PH_DEFAULT, RPL>(a, grid, s); break;
    }
}

void launch_fwd(const FwdArgs& a, int variant, int rpl, int grid, hipStream_t s) {
    if (rpl == 2) launch_fwd_r<2>(a, variant, grid, s);
    else launch_fwd_r<1>(a, variant, grid, s);
}

static inline dim3 g1(size_t n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

void launch_lin(const LinArgs& a, int variant, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((a.sc.P + 255) / 256), static_cast<unsigned>(a.nsteps));
    switch (variant) {
        case PH_DDR: hipLaunchKernelGGL(k_adj_lin<PH_DDR>, grid, dim3(256), 0, s, a); break;
        case PH_POS: hipLaunchKernelGGL(k_adj_lin<PH_POS>, grid, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL(k_adj_lin<PH_DEFAULT>, grid, dim3(256), 0, s, a); break;
    }
}
void launch_sweep(const SweepArgs& a, int rpl, int grid, hipStream_t s) {
    if (rpl == 2) hipLaunchKernelGGL(k_adj_sweep<2>, dim3(grid), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_adj_sweep<1>, dim3(grid), dim3(64), 0, s, a);
}
void launch_fin(const FinArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_fin, g1(a.P), dim3(256), 0, s, a);
}
void launch_tri_fwd(const TriArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_fwd, dim3(grid), dim3(64), 0, s, a);
}
void launch_tri_adj(const TriArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_adj, dim3(grid), dim3(64), 0, s, a);
}

void launch_grad_a(int nnz, const int* row, const int* col, const float* gb, const float* x,
                   float* ga, hipStream_t s) {
    hipLaunchKernelGGL(k_grad_a, g1(nnz), dim3(256), 0, s, nnz, row, col, gb, x, ga);
}
void launch_bind_statics(const Sched& sc, const float* n, const float* qsp, const float* psp,
                         const float* len, const float* slope, const float* xst, float slope_lb,
                         float4* st, hipStream_t s) {
    hipLaunchKernelGGL(k_bind_statics, g1(sc.P), dim3(256), 0, s, sc, n, qsp, psp, len, slope, xst,
                       slope_lb, st);
}
void launch_state(const Sched& sc, const float* q, int clamp, float lb, float* st_q,
                  float* st_i, float* Q_row0, hipStream_t s) {
    hipLaunchKernelGGL(k_state_q, g1(sc.P), dim3(256), 0, s, sc, q, clamp, lb, st_q, Q_row0);
    hipLaunchKernelGGL(k_state_i, g1(sc.P), dim3(256), 0, s, sc, st_q, st_i);
}
void launch_hist_from_x(const Sched& sc, const float* x, uint64_t* hist, hipStream_t s) {
    hipLaunchKernelGGL(k_hist_from_x, g1(sc.P), dim3(256), 0, s, sc, x, hist);
}
void launch_lam_init(const Sched& sc, const float* g, const uint64_t* hist_row, float lb,
                     float* lam, float* a0, float* a1, float* a2, hipStream_t s) {
    hipLaunchKernelGGL(k_lam_init, g1(sc.P), dim3(256), 0, s, sc, g, hist_row, lb, lam, a0, a1,
                       a2);
}
void launch_unpermute(const Sched& sc, const float* src, float* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_unpermute, g1(sc.P), dim3(256), 0, s, sc, src, dst);
}
void launch_grad_q0(const Sched& sc, const float* lam, const float* q0, float lb, float* g,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_grad_q0, g1(sc.P), dim3(256), 0, s, sc, lam, q0, lb, g);
}
void launch_zero(size_t n, float* x, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_zero_f, g1(n), dim3(256), 0, s, n, x);
}

template <class K>
static int occ(K kernel) {
    int nb = 0;
    return (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64, 0) == hipSuccess) ? nb : 0;
}

int occupancy_blocks(int kernel, int variant, int rpl) {
    switch (kernel) {
        case OCC_FWD:
            if (rpl == 2) {
                if (variant == PH_DDR) return occ(k_fwd_window<PH_DDR, 2>);
                if (variant == PH_POS) return occ(k_fwd_window<PH_POS, 2>);
                return occ(k_fwd_window<PH_DEFAULT, 2>);
            }
            if (variant == PH_DDR) return occ(k_fwd_window<PH_DDR, 1>);
            if (variant == PH_POS) return occ(k_fwd_window<PH_POS, 1>);
            return occ(k_fwd_window<PH_DEFAULT, 1>);
        case OCC_SWEEP: return (rpl == 2) ? occ(k_adj_sweep<2>) : occ(k_adj_sweep<1>);
        case OCC_TRI_FWD: return occ(k_tri_fwd);
        default: return occ(k_tri_adj);
    }
}

}  // namespace ddrs

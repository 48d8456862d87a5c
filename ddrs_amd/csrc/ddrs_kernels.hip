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

#include <algorithm>

#include "mc_physics.h"
#include "ddrs_epoch.h"
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

// Pairs of tagged words: two words of 16 bytes, 16-byte aligned, moved by one
// vector access with the cache policy of the agent-scope atomics (sc1, device
// coherent). Each 8-byte half keeps its own tag, so a reader that takes a pair
// only when both tags are current never mixes two publications, even if the
// 16 bytes were moved as two 8-byte halves. The buffer resource covers one
// step's row of pairs (uniform, in scalar registers); the offset is in bytes.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCoherentPolicy = 16;  // sc1
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pair_row(const uint64_t* row, int npairs) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(row), 0, npairs * 16,
                                             0x00020000);
}
__device__ __forceinline__ void publish_pair(__amdgpu_buffer_rsrc_t row, int i, float a, float b,
                                             uint32_t epoch) {
    const u32x4 v = {__float_as_uint(a), epoch, __float_as_uint(b), epoch};
    __builtin_amdgcn_raw_buffer_store_b128(v, row, i * 16, 0, kCoherentPolicy);
}
__device__ __forceinline__ u32x4 poll_pair(__amdgpu_buffer_rsrc_t row, int i) {
    return __builtin_amdgcn_raw_buffer_load_b128(row, i * 16, 0, kCoherentPolicy);
}
__device__ __forceinline__ bool pair_ready(const u32x4& w, uint32_t tag) {
    return w.y == tag && w.w == tag;
}

// One row of an array (`bytes` long, at most 2 GiB) as a buffer resource: its
// base stays in scalar registers and a lane addresses it with a 32-bit byte
// offset, so a persistent kernel holds no 64-bit address per lane and array
// across its steps (the compiler keeps such addresses live across the whole
// loop and spills them). Accesses past `bytes` read 0 and write nothing.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes),
                                             0x00020000);
}
// tagged words through a row: the cache policy of the agent-scope atomics
__device__ __forceinline__ uint64_t poll_row(__amdgpu_buffer_rsrc_t row, int i) {
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(row, i * 8, 0, kCoherentPolicy);
    return (static_cast<uint64_t>(w.y) << 32) | w.x;
}
__device__ __forceinline__ void publish_row(__amdgpu_buffer_rsrc_t row, int i, float v,
                                            uint32_t epoch) {
    const u32x2 w = {__float_as_uint(v), epoch};
    __builtin_amdgcn_raw_buffer_store_b64(w, row, i * 8, 0, kCoherentPolicy);
}
__device__ __forceinline__ float load_row(__amdgpu_buffer_rsrc_t row, int i) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(row, i * 4, 0, 0));
}
__device__ __forceinline__ void store_row(__amdgpu_buffer_rsrc_t row, int i, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), row, i * 4, 0, 0);
}
__device__ __forceinline__ void store_row2(__amdgpu_buffer_rsrc_t row, int i, float a, float b) {
    const u32x2 w = {__float_as_uint(a), __float_as_uint(b)};
    __builtin_amdgcn_raw_buffer_store_b64(w, row, i * 8, 0, 0);
}

// Waits until at most N vector memory operations of the wave are outstanding.
// They complete in issue order, so this waits for everything issued before
// the last N.
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// LDS reads of data that direct-to-LDS loads wrote, for use after wait_vm
// has covered those loads. The compiler orders its own LDS reads after every
// outstanding direct-to-LDS load by waiting for all vector memory operations,
// stores included; these reads carry their own LDS wait and nothing else.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p));
}
__device__ __forceinline__ u32x4 lds_read_dma16(const void* p) {
    u32x4 v;
    asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_addr(p)) : "memory");
    return v;
}
__device__ __forceinline__ float lds_read_dma4(const void* p) {
    float v;
    asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_addr(p)) : "memory");
    return v;
}

// Orders LDS within a one-wave workgroup: a wave's LDS operations execute in
// issue order, so a compiler barrier is all it needs (__syncthreads would also
// wait for every outstanding global store, to release them to the workgroup).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Loads `bytes` (4 or 16) per lane from g straight into LDS at lds + lane *
// bytes (lds the same for the whole wave); the data passes through no
// register. The compiler waits for the load before LDS reads of that data.
__device__ __forceinline__ void load_to_lds(const void* g, void* lds, int bytes) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (bytes == 16) __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
    else __builtin_amdgcn_global_load_lds(g, lds, 4, 0, 0);
#endif
}
// The same for 16 bytes with the cache policy of the agent-scope atomics
// (sc1, device coherent): polls of tagged pairs.
__device__ __forceinline__ void poll_pair_to_lds(const void* g, void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, kCoherentPolicy);
#endif
}

// 100 MHz wall clock
__device__ __forceinline__ uint64_t wall_ticks() { return __builtin_amdgcn_s_memrealtime(); }
constexpr uint64_t kWatchdogTicks = 200000000ull;  // 2 s at 100 MHz
constexpr uint64_t kErrCheckTicks = 2000ull;        // 20 us at 100 MHz

// ---------------------------------------------------------------------------
// One chunk per workgroup, for all steps of the launch.

// Chunk of this workgroup: the claim counter only grows (claims of earlier
// launches are subtracted through the launch's claim base), so it is never
// reset. Dir = +1 hands out chunks 0, 1, ..., Dir = -1 hands out C-1, C-2, ...
// A claim outside [0, C) (a launch whose state was not begun) returns -1 with
// the error flag raised; the caller's workgroup then returns at once.
template <int Dir>
__device__ __forceinline__ int claim_chunk(const Claim& cl, int C, int* lds, int* err) {
    if (threadIdx.x == 0) {
        const unsigned long long k = atomicAdd(cl.ctr, 1ull) - cl.ls->claim_base;
        const bool ok = k < static_cast<unsigned long long>(C);
        if (!ok) atomicExch(err, 1);
        *lds = ok ? static_cast<int>(k) : -1;
    }
    __syncthreads();
    const int k = uni(*lds);
    if (k < 0) return -1;
    return (Dir > 0) ? k : (C - 1 - k);
}

// The tag of the words this launch publishes (see LaunchState).
__device__ __forceinline__ uint32_t launch_epoch(const Claim& cl) { return cl.ls->epoch; }

__global__ void k_begin_launch(LaunchState* ls, const unsigned long long* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ls->claim_base = *ctr;
        ls->epoch = next_epoch(ls->epoch);
    }
}

// For the launch's steps in time order (Dir = +1) or reverse (Dir = -1):
//   issue(s, again)    starts the loads of the cross-chunk inputs of step s
//                      (device-coherent, so each costs a full round trip);
//                      `again` when the inputs are read again while waiting
//                      (anything else issue(s) loads need not be reloaded);
//   prepare(s, first)  everything of step s that does not depend on other
//                      chunks (coefficients, right-hand side, into LDS);
//                      `first` asks for the chunk's step-invariant data too;
//   check(s)           waits for the issued loads and stages the inputs in
//                      LDS; true when all of them belong to step s;
//   probe(s)           while waiting: polls one representative input per
//                      producing chunk (the one it publishes last); true when
//                      all of them belong to step s, so the inputs are read again;
//   solve(s, last, h)  the level-by-level solve inside the chunk, publishing,
//                      and the step's epilogue; `last` is the launch's final step.
//                      It calls h() after its arithmetic and before its first
//                      store: h issues the next step's loads.
// The next step's inputs are requested before this step's stores: the wave's
// memory counter is in order, so a wait for loads issued after the stores
// would also wait for the stores to be acknowledged (a round trip to the
// coherence point), while loads issued before them are waited for alone. The
// requests also overlap the next step's coefficient arithmetic. A waiting
// chunk polls only the representatives, after a growing pause (0.2 to 0.8 us),
// so that waiting chunks do not flood the memory system. All five are called
// by the whole wave, and the workgroup is that one wave: its LDS operations
// execute in order, so the phases need no workgroup barrier. Returns early
// (error flag set) when the inputs of a step do not arrive within the watchdog
// time or another workgroup has raised the flag.
template <int Dir, class Issue, class Prepare, class Check, class Probe, class Solve>
__device__ __forceinline__ void chunk_loop(int nsteps, int* err, unsigned long long* trace,
                                           Issue issue, Prepare prepare, Check check,
                                           Probe probe, Solve solve, bool raised = false) {
    const uint64_t t_claim = wall_ticks();
    // Issue priority outside the solve: raised (2) for chunks the caller marks,
    // e.g. deep chunks of the forward, whose steps the rest of the window
    // waits on, while shallow chunks run ahead with slack; the solve runs at 3.
    auto base_prio = [&]() {
        if (raised) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
    };
    base_prio();
    uint64_t t_first = 0, t_prep = 0, t_poll = 0, t_wait = 0, t_solve = 0;
    unsigned long long missed = 0;
    for (int done = 0; done < nsteps; ++done) {
        const int s = (Dir > 0) ? done : (nsteps - 1 - done);
        const uint64_t ta = trace ? wall_ticks() : 0;
        prepare(s, done == 0);
        wave_lds_sync();
        if (done == 0) issue(s, false);  // the first step stages its links in prepare
        const uint64_t tb = trace ? wall_ticks() : 0;
        const bool ok = check(s);
        const uint64_t tc = trace ? wall_ticks() : 0;
        if (!ok) {
            const uint64_t t0 = wall_ticks();
            int nap = 1;
            for (;;) {
                ++missed;
                for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(8);
                nap = (nap < 4) ? 2 * nap : 4;
                // The error flag is a device-coherent read, a full memory round
                // trip: a wait reads it only once it has lasted kErrCheckTicks,
                // so that the usual short waits spend their round trips on
                // their inputs alone.
                const uint64_t waited = wall_ticks() - t0;
                const bool give_up =
                    waited > kErrCheckTicks &&
                    (waited > kWatchdogTicks ||
                     __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0);
                if (give_up) {
                    if (threadIdx.x == 0) atomicExch(err, 1);
                    return;
                }
                if (!probe(s)) continue;
                issue(s, true);
                if (check(s)) break;
            }
        }
        const uint64_t td = trace ? wall_ticks() : 0;
        // The solve is a chain of dependent LDS round trips; it issues ahead of
        // the other waves on the SIMD, whose arithmetic fills the gaps.
        __builtin_amdgcn_s_setprio(3);
        auto issue_next = [&]() {
            if (done + 1 < nsteps) issue((Dir > 0) ? s + 1 : s - 1, false);
        };
        solve(s, done == nsteps - 1, issue_next);
        base_prio();
        if (trace) {
            const uint64_t te = wall_ticks();
            if (done == 0) t_first = te;
            t_prep += tb - ta;
            t_poll += tc - tb;
            t_wait += td - tc;
            t_solve += te - td;
        }
    }
    if (trace && threadIdx.x == 0) {
        trace[0] = t_claim;
        trace[1] = t_first;
        trace[2] = wall_ticks();
        trace[3] = missed;
        trace[4] = t_prep;
        trace[5] = t_poll;
        trace[6] = t_wait;
        trace[7] = t_solve;
    }
}

// (v / ad) for the diagonal ad = 1 + (-c1) * 0 of I - c1 N: exactly v when
// ad == 1, the division otherwise (c1 infinite or NaN).
__device__ __forceinline__ float div_diag(float v, float ad) { return (ad == 1.0f) ? v : v / ad; }


template <int RPL>
struct Width {
    static constexpr int CW = 64 * RPL;                  // positions per chunk
    static constexpr int SHIFT = (RPL == 1) ? 6 : 7;
    static constexpr int kStageUp = kStageUpBase * RPL;  // staged upstream links
    static constexpr int kStageDn = kStageDnBase * RPL;  // staged downstream links
    // forward: link slots per lane whose polls are issued ahead of prepare
    // (every staged link), and upstream links per reach held in registers
    static constexpr int kSlots = kStageUp / 64;
    static constexpr int kFast = 4;
};

// A reach's first links, kept per lane so that a sub-level of the solve reads
// only the linked values. Adjoint: code >= 0 is the chunk-local index of a
// link inside the chunk, code < 0 is -1 - (its slot among the chunk's staged
// links). Forward: one index into the chunk's values, [0, CW) the chunk's own
// solves and CW + slot the staged links from other chunks, so that every link
// of a sub-level is one LDS read, all issued before the first is waited for.
// The first Width::kFast upstream links of a reach are read that way (four
// cover all but ~0.4% of the reaches of river networks; the register budget
// of six waves per SIMD holds no more); the links beyond them take a loop.
constexpr int kFastLinksDn = 1;  // adjoint: downstream links (one per reach in a tree)
// Sub-levels whose largest link count is kept per chunk (4 bits each).
constexpr int kLinkBoundSubs = 16;
template <int RPL>
__device__ __forceinline__ int link_code(int c, int base, int e, int q) {
    return ((q >> Width<RPL>::SHIFT) == c) ? (q & (Width<RPL>::CW - 1)) : (-1 - (e - base));
}
template <int RPL>
__device__ __forceinline__ int link_index(int c, int base, int e, int q) {
    return ((q >> Width<RPL>::SHIFT) == c) ? (q & (Width<RPL>::CW - 1))
                                          : (Width<RPL>::CW + (e - base));
}

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
__global__ __launch_bounds__(64, 6) void k_fwd_window(FwdArgs a) {
    using Wd = Width<RPL>;
    constexpr int CW = Wd::CW;
    // the chunk's solves (lds_x) and the staged values of links from other
    // chunks (sval), one array so that a link is one index (link_index)
    __shared__ float xv[CW + Wd::kStageUp];
    float* const lds_x = xv;
    float* const sval = xv + CW;
    __shared__ float sadj[Wd::kStageUp];
    __shared__ int spos[Wd::kStageUp];
    __shared__ short skind[Wd::kStageUp];
    __shared__ float pb[CW], pnc1[CW], sq[CW], si[CW];
    __shared__ short pe0[CW], pe1[CW], psub[CW];
    __shared__ int porig[CW];
    __shared__ int pgs[CW];  // gauge slot of the position, -1: none
    __shared__ short srep[kRepMax];
    __shared__ short lkc[Wd::kFast][CW];  // link_index (< CW + kStageUp)
    __shared__ int slm[kLinkBoundSubs];
    __shared__ LnEnt sln[kLnEntries];  // the logarithm table of the powers
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int R = sc.R, P = sc.P;
    const int c = claim_chunk<+1>(a.claim, a.ch.C, &claimed, a.err);
    if (c < 0) return;
    const uint32_t epoch = launch_epoch(a.claim);
    bool unit = true;  // every diagonal of this step's I - c1 N is exactly 1
    const int rep0 = a.ch.rep_up_ptr[c];
    const int nrep_all = a.ch.rep_up_ptr[c + 1] - rep0;
    const int nrep = (nrep_all <= kRepMax) ? nrep_all : -1;
    const int p0 = c * CW;
    const int base = sc.up_ptr[p0];
    const int nlinks = sc.up_ptr[p0 + CW] - base;
    const int nsub = a.ch.nsub[c];

    // Reach statics stay in registers for the whole launch; q' of the next
    // step is loaded one step ahead, straight into LDS (sqp): a register
    // carrying the load into the next step would make the compiler wait, in
    // the middle of the coefficient arithmetic, for everything in flight (the
    // previous step's stores included) to move it into place. check() waits
    // for all loads, so the next prepare reads sqp without a wait of its own.
    Reach rs[RPL];
    __shared__ float sqp[CW];
    // whether this chunk holds a gauge of the readout (wave-uniform)
    bool hg = false;
    // links to read at sub-level l < kLinkBoundSubs: bits [4l, 4l + 4) hold the
    // largest link count (capped at kFast) of the sub-level's reaches; lall
    // the largest over the chunk (wave-uniform)
    uint64_t lbound = ~0ull;
    int lall = Wd::kFast;
    // link slots k this lane polls (bit k: staged link lane + 64 k is from
    // another chunk), so that no step reads the staging tables to decide
    unsigned pollm = 0;
    int rpos = 0;  // the position this lane probes (lane < nrep)

    // The chunk's step-invariant data, before the first step: statics into
    // registers, links, gauge slots, positions and per-sub-level link bounds
    // into LDS, and q' of the first step (sqp). Every load is waited for here,
    // so that no register is still loading when the step loop starts (the
    // compiler would otherwise wait, at its uses in every step, for every
    // operation in flight), and the pointers used only here are dead inside
    // the loop.
    {
        const int t = a.t0 + a.s_base;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int p = p0 + r * 64 + lane;
            rs[r] = load_reach(a.st, p);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                row_rsrc(a.qps + static_cast<size_t>(t) * P, P * 4ull), &sqp[r * 64], 4, p * 4,
                0, 0, 0);
        }
        stage_links<RPL>(c, base, base + nlinks, sc.up_pos, sc.up_adj, spos, sadj, skind);
        if (lane < nrep) srep[lane] = a.ch.rep_up[rep0 + lane];
        if (a.Qg) {
            int any = 0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int sl = a.gslot[p0 + r * 64 + lane];
                pgs[r * 64 + lane] = sl;
                any |= sl >= 0;
            }
            hg = __any(any) != 0;
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int p = p0 + j;
            const int orig = sc.perm[p];
            porig[j] = orig;
            psub[j] = -1;
            if (orig >= 0) {
                const int ea = sc.up_ptr[p], eb = sc.up_ptr[p + 1];
                pe0[j] = static_cast<short>(ea - base);
                pe1[j] = static_cast<short>(eb - base);
                psub[j] = a.ch.sub[p];
                sq[j] = a.st_q[p];
                si[j] = a.st_i[p];
#pragma unroll
                for (int k = 0; k < Wd::kFast; ++k) {
                    if (ea + k < eb)
                        lkc[k][j] = static_cast<short>(
                            link_index<RPL>(c, base, ea + k, sc.up_pos[ea + k]));
                }
            }
        }
        if (lane < kLnEntries) sln[lane] = kLnTab[lane];
        wait_vm<0>();
        // per sub-level link bounds, for the solve's read count
        if (lane < kLinkBoundSubs) slm[lane] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int l = psub[j];
            if (l >= 0 && l < kLinkBoundSubs) atomicMax(&slm[l], min(pe1[j] - pe0[j], Wd::kFast));
        }
        __syncthreads();
        uint64_t m = 0;
        int mx = 0;
        for (int l = 0; l < kLinkBoundSubs; ++l) {
            const int v = slm[l];
            m |= static_cast<uint64_t>(v) << (4 * l);
            mx = max(mx, v);
        }
        lbound = (static_cast<uint64_t>(uni(static_cast<int>(m >> 32))) << 32) |
                 static_cast<uint32_t>(uni(static_cast<int>(m)));
        // sub-levels beyond the kept ones read every register-held link
        lall = (nsub > kLinkBoundSubs) ? Wd::kFast : uni(mx);
#pragma unroll
        for (int k = 0; k < Wd::kSlots; ++k) {
            const int e = lane + 64 * k;
            if (e < nlinks && skind[e] < 0) pollm |= 1u << k;
        }
        if (lane < nrep) rpos = spos[srep[lane]];
    }

    // S1..S25 for window step s: coefficients and right-hand side.
    auto prepare = [&](int s, bool) {
        bool ok_diag = true;
        const int t = a.t0 + a.s_base + s;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int p = p0 + j;
            if (porig[j] >= 0) {
                const float q = sq[j];
                const float it = si[j];
                const float qp = fmaxr(lds_read_dma4(&sqp[j]), a.P.q_lb);
                if (s + 1 < a.nsteps)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        row_rsrc(a.qps + static_cast<size_t>(t + 1) * P, P * 4ull), &sqp[r * 64], 4,
                        p * 4, 0, 0, 0);
                Coef co;
                PowEvalInl pe(sln);
                mc_coef<PH, PowEvalInl, true>(a.P, rs[r], q, co, pe);
                pb[j] = ((co.c2 * it) + (co.c3 * q)) + (co.c4 * qp);             // S25
                pnc1[j] = -co.c1;
                ok_diag = ok_diag && (1.0f + (-co.c1) * 0.0f == 1.0f);
            }
        }
        unit = __all(ok_diag);
    };

    // every staged link is polled before prepare, so that the check after it
    // waits for no round trip when the inputs were published in time
    uint64_t pw[Wd::kSlots];
    auto issue = [&](int s, bool) {
        const auto row = row_rsrc(a.hist + static_cast<size_t>(a.s_base + s) * P, P * 8ull);
#pragma unroll
        for (int k = 0; k < Wd::kSlots; ++k) {
            if (64 * k < nlinks) {
                const int e = lane + 64 * k;
                pw[k] = ((pollm >> k) & 1u) ? poll_row(row, spos[e]) : 0;
            }
        }
    };
    auto check = [&](int) -> bool {
        // every load in flight, q' of the next step (sqp) included
        wait_vm<0>();
        bool ok = true;
#pragma unroll
        for (int k = 0; k < Wd::kSlots; ++k) {
            if (64 * k < nlinks) {
                const int e = lane + 64 * k;
                if ((pollm >> k) & 1u) {
                    if (!word_ready(pw[k], epoch)) ok = false;
                    else sval[e] = word_value(pw[k]);
                }
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };

    auto probe = [&](int s) -> bool {
        if (nrep <= 0) return true;  // no representatives: read the links again
        const auto row = row_rsrc(a.hist + static_cast<size_t>(a.s_base + s) * P, P * 8ull);
        bool ok = true;
        if (lane < nrep) ok = word_ready(poll_row(row, rpos), epoch);
        return __all(ok);
    };

    // S26..S28 and the next step's S24 (N Q_{t+1}) from LDS only.
    auto solve = [&](int s, bool last, auto&& before_stores) {
        const int sg = a.s_base + s;
        const int t = a.t0 + sg;
        const auto row = row_rsrc(a.hist + static_cast<size_t>(sg) * P, P * 8ull);
        // per-lane operands of the solve, read once per step
        // link indices two to a register (16 bits each); the link weights are
        // read beside the linked values from the staged weights (slot e0 + k)
        int my[RPL], e0[RPL], nl[RPL];
        uint32_t lix2[RPL][(Wd::kFast + 1) / 2];
        float x[RPL], inext[RPL], sxn[RPL], b[RPL], nc1[RPL];
        auto lix = [&](int r, int k) -> int {
            return static_cast<int>((lix2[r][k >> 1] >> (16 * (k & 1))) & 0xffffu);
        };
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            my[r] = psub[j];
            x[r] = 0.0f;
            inext[r] = 0.0f;
            sxn[r] = 0.0f;
            b[r] = pb[j];
            nc1[r] = pnc1[j];
            e0[r] = pe0[j];
            nl[r] = pe1[j] - e0[r];
#pragma unroll
            for (int k = 0; k < (Wd::kFast + 1) / 2; ++k) {
                // unused slots read value 0 of the chunk (always in bounds)
                const uint32_t i0 = (2 * k < nl[r]) ? static_cast<uint32_t>(lkc[2 * k][j]) : 0u;
                const uint32_t i1 = (2 * k + 1 < Wd::kFast && 2 * k + 1 < nl[r])
                                        ? static_cast<uint32_t>(lkc[2 * k + 1][j]) : 0u;
                lix2[r][k] = i0 | (i1 << 16);
            }
        }
        // The sub-levels carry only the solve's own sum; the step's other two
        // upstream sums (N max(x, lb) for the next right-hand side and N x for
        // the history) are taken once after the last sub-level, over the same
        // values in the same CSR order.
        for (int l = 0; l < nsub; ++l) {
            const int lm = (l < kLinkBoundSubs) ? static_cast<int>((lbound >> (4 * l)) & 15u)
                                                : Wd::kFast;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (my[r] == l) {
                    const int j = r * 64 + lane;
                    float sum = 0.0f;
                    {
                        // the first kFast links: every read first, then the sum
                        // in CSR order
                        float v[Wd::kFast];
                        float wk[Wd::kFast];
#pragma unroll
                        for (int k = 0; k < Wd::kFast; ++k) {
                            if (k < lm) {
                                v[k] = xv[lix(r, k)];
                                wk[k] = sadj[(k < nl[r]) ? e0[r] + k : 0];
                            }
                        }
#pragma unroll
                        for (int k = 0; k < Wd::kFast; ++k) {
                            if (k < lm) {
                                const float term = (nc1[r] * wk[k]) * v[k];
                                if (k < nl[r]) sum = sum + term;
                            }
                        }
                    }
                    // links beyond them (rare), continuing the order
                    for (int e = e0[r] + Wd::kFast; e < e0[r] + nl[r]; ++e) {
                        const int kd = skind[e];
                        const float v = (kd >= 0) ? lds_x[kd] : sval[e];
                        sum = sum + (nc1[r] * sadj[e]) * v;
                    }
                    x[r] = unit ? (b[r] - sum) : div_diag(b[r] - sum, 1.0f + nc1[r] * 0.0f);
                    lds_x[j] = x[r];
                }
            }
            __syncthreads();
        }
        before_stores();
        // Publish as soon as the solve is done (other chunks read a step's
        // values only then), before the step's two upstream sums.
#pragma unroll
        for (int r = 0; r < RPL; ++r)
            if (my[r] >= 0) publish_row(row, p0 + r * 64 + lane, x[r], epoch);
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            float in = 0.0f, sx = 0.0f;
            if (my[r] >= 0) {
                float v[Wd::kFast], wk[Wd::kFast];
#pragma unroll
                for (int k = 0; k < Wd::kFast; ++k) {
                    if (k < lall) {
                        v[k] = xv[lix(r, k)];
                        wk[k] = sadj[(k < nl[r]) ? e0[r] + k : 0];
                    }
                }
#pragma unroll
                for (int k = 0; k < Wd::kFast; ++k) {
                    if (k < lall) {
                        const float ti = fmaxr(v[k], a.P.q_lb) * wk[k];
                        const float tx = v[k] * wk[k];
                        if (k < nl[r]) {
                            in = in + ti;
                            sx = sx + tx;
                        }
                    }
                }
                for (int e = e0[r] + Wd::kFast; e < e0[r] + nl[r]; ++e) {
                    const int kd = skind[e];
                    const float vv = (kd >= 0) ? lds_x[kd] : sval[e];
                    const float adj = sadj[e];
                    in = in + fmaxr(vv, a.P.q_lb) * adj;
                    sx = sx + vv * adj;
                }
            }
            inext[r] = in;
            sxn[r] = sx;
        }
        if (a.ihist) {
            const auto irow = row_rsrc(a.ihist + static_cast<size_t>(sg) * P, P * 8ull);
#pragma unroll
            for (int r = 0; r < RPL; ++r)
                if (my[r] >= 0) store_row2(irow, p0 + r * 64 + lane, sxn[r], inext[r]);
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
                if (a.Q_out)
                    store_row(row_rsrc(a.Q_out + static_cast<size_t>(t + 1) * R, R * 4ull), porig[j], qn);
                if (a.X_out) a.X_out[porig[j]] = x[r];
                if (hg) {
                    const int sl = pgs[j];
                    if (sl >= 0)
                        store_row(row_rsrc(a.Qg + static_cast<size_t>(t + 1) * a.gK, a.gK * 4ull), sl, qn);
                }
            }
        }
    };

    chunk_loop<+1>(a.nsteps, a.err, a.trace.rec ? a.trace.rec + kTraceWords * c : nullptr, issue,
                   prepare, check, probe, solve, c >= a.prio_from);
}

// ---------------------------------------------------------------------------
// adjoint window, phase 1: linearisation of every reach-step
//
// One step's backward (B28..B1) is linear in the transposed-solve value gb once
// the forward state is fixed. This kernel evaluates, for every reach and step
// of the segment, the coefficients (S1..S23 from the forward history) and the
// backward chain for gb = 1; the reverse-time sweep then only solves for gb,
// and the parameter gradients are gb times these records. A thread walks one
// reach through kLinSteps consecutive steps: its statics and links are loaded
// once, and each history word (own and upstream x of a step) is read once and
// carried into the next step as out[t] and N out[t]. The backward chain scales
// gradients with the hardware reciprocal (about 1 ulp); the coefficients and
// every clamp decision are computed exactly as in the forward.

constexpr int kLinSteps = 32;
constexpr int kLinBlock = 256;   // positions per workgroup of the one-launch grid
constexpr int kLinItem = 64;     // positions per item of a launch on the item counter
// A concurrent launch waits for the forward's claims while they advance; it
// leaves when they stall this long (its own waves may be what keeps the
// forward's last workgroups from becoming resident).
constexpr uint64_t kClaimStallTicks = 5000;  // 50 us at 100 MHz

// Steps [s_first, s_first + kLinSteps) of position p. Coh: the forward that
// writes the history may still run; history words are read with coherent
// loads and checked for its epoch (awaited when missing).
template <int PH, bool Coh>
__device__ __forceinline__ void lin_item(const LinArgs& a, int sb, int p, const LnEnt* lnt) {
    const Sched& sc = a.sc;
    const int P = sc.P;
    if (p >= P) return;
    const int orig = sc.perm[p];
    if (orig < 0) return;
    const int s_first = a.s_base + sb * kLinSteps;
    const int s_end = min(s_first + kLinSteps, a.s_base + a.nsteps);
    const float lb = a.P.q_lb;
    const int e0 = sc.up_ptr[p], e1 = sc.up_ptr[p + 1];
    const Reach rc = load_reach(a.st, p);
    const uint32_t ep = a.fwd_epoch;
    // history word, its epoch folded into `ok`
    auto rd = [&](const uint64_t* w, bool& ok) -> float {
        if constexpr (!Coh) {
            return word_value(*w);
        } else {
            const uint64_t v = poll(w);
            ok = ok && word_ready(v, ep);
            return word_value(v);
        }
    };
    // the same word, awaited until the forward has written it; gives up (error
    // flag set) after the watchdog time or once any workgroup has raised the
    // flag, as the forward's workgroups do when they give up
    auto wt = [&](const uint64_t* w) -> float {
        uint64_t v = poll(w);
        const uint64_t t0 = wall_ticks();
        int nap = 1;
        while (!word_ready(v, ep)) {
            for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(8);
            nap = (nap < 8) ? 2 * nap : 8;
            if (wall_ticks() - t0 > kWatchdogTicks ||
                __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                atomicExch(a.err, 1);
                return 0.0f;
            }
            v = poll(w);
        }
        return word_value(v);
    };
    // sx = sum_k adj_k x[up k] (B26 at gb = 1: gc1 = -sum_k adj_k ((-1) x[col k]))
    // and in = sum_k adj_k max(x[up k], lb) = N out over the upstream of p in `row`
    auto inflow = [&](const uint64_t* row, float& sx, float& in, bool& ok) {
        sx = 0.0f;
        in = 0.0f;
        for (int e = e0; e < e1; ++e) {
            const float xu = rd(row + sc.up_pos[e], ok);
            const float w = sc.up_adj[e];
            sx = sx + xu * w;
            in = in + fmaxr(xu, lb) * w;
        }
    };
    auto inflow_wait = [&](const uint64_t* row, float& sx, float& in) {
        sx = 0.0f;
        in = 0.0f;
        for (int e = e0; e < e1; ++e) {
            const float xu = wt(row + sc.up_pos[e]);
            const float w = sc.up_adj[e];
            sx = sx + xu * w;
            in = in + fmaxr(xu, lb) * w;
        }
    };
    if constexpr (Coh) {
        // start once this position's last step of the item is out; its
        // upstream words were published before it
        (void)wt(a.hist + static_cast<size_t>(s_end - 1) * P + p);
    }
    // state entering the first step: out[t] and N out[t]
    float q, it;
    const bool have_sums = !Coh && a.ihist != nullptr;
    if (s_first > 0) {
        const uint64_t* hq = a.hist + static_cast<size_t>(s_first - 1) * P;
        bool ok = true;
        float xq = rd(hq + p, ok), sx_unused;
        if (have_sums) it = a.ihist[static_cast<size_t>(s_first - 1) * P + p].y;
        else inflow(hq, sx_unused, it, ok);
        if (Coh && !ok) {
            xq = wt(hq + p);
            inflow_wait(hq, sx_unused, it);
        }
        q = fmaxr(xq, lb);
    } else {
        q = a.ck_q[p];
        it = a.ck_i[p];
    }
    for (int sg = s_first; sg < s_end; ++sg) {
        const int t = a.t0 + sg;
        const uint64_t* hx = a.hist + static_cast<size_t>(sg) * P;
        bool ok = true;
        float sx, inext;
        if (have_sums) {
            const float2 u = a.ihist[static_cast<size_t>(sg) * P + p];
            sx = u.x;
            inext = u.y;
        } else {
            inflow(hx, sx, inext, ok);
        }
        float xs = rd(hx + p, ok);
        if (Coh && !ok) {
            inflow_wait(hx, sx, inext);
            xs = wt(hx + p);
        }
        const float qp_raw = a.qps[static_cast<size_t>(t) * P + p];
        const float qp = fmaxr(qp_raw, lb);
        Coef co;
        PowEvalInl pe(lnt);
        mc_coef<PH, PowEvalInl, true>(a.P, rc, q, co, pe);
        const LocalGrads g = mc_local_bwd<PH, true>(a.P, co, rc, q, qp, it, 1.0f, sx, 0.0f);
        const size_t o = static_cast<size_t>(sg) * P + p;
        a.rec[o] = make_float4(-co.c1, co.c2, g.gq, (!a.mask_lam || q > lb) ? 1.0f : 0.0f);
        a.recj[o] = g.gn;
        a.recj[a.plane + o] = g.gqsp;
        a.recj[2 * a.plane + o] = g.gp;
        a.recj[3 * a.plane + o] = (!a.mask_qp || qp_raw > lb) ? g.gqp : 0.0f;
        q = fmaxr(xs, lb);
        it = inext;
    }
}

// One workgroup per item: grid x = block of kLinBlock positions, y = step block.
template <int PH>
__global__ __launch_bounds__(kLinBlock, 6) void k_adj_lin(LinArgs a) {
    __shared__ LnEnt sln[kLnEntries];  // the logarithm table of the powers
    if (threadIdx.x < kLnEntries) sln[threadIdx.x] = kLnTab[threadIdx.x];
    __syncthreads();
    lin_item<PH, false>(a, static_cast<int>(blockIdx.y),
                        static_cast<int>(blockIdx.x) * kLinBlock + static_cast<int>(threadIdx.x), sln);
}

// One-wave workgroups taking items of kLinItem positions from the counter, in
// step-block-major order.
template <int PH, bool Coh>
__global__ __launch_bounds__(kLinItem) void k_adj_lin_items(LinArgs a) {
    const int npb = (a.sc.P + kLinItem - 1) / kLinItem;
    __shared__ LnEnt sln[kLnEntries];  // the logarithm table of the powers
    if (threadIdx.x < kLnEntries) sln[threadIdx.x] = kLnTab[threadIdx.x];
    __syncthreads();
    if constexpr (Coh) {
        // wait for the forward's claims while they advance
        const auto* ctr = const_cast<unsigned long long*>(a.fwd_claim_ctr);
        unsigned long long seen = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t t_seen = wall_ticks();
        while (seen < a.fwd_claims) {
            __builtin_amdgcn_s_sleep(8);
            const unsigned long long v =
                __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t now = wall_ticks();
            if (v != seen) {
                seen = v;
                t_seen = now;
            } else if (now - t_seen > kClaimStallTicks) {
                if (a.trace && threadIdx.x == 0) atomicAdd(a.trace + kTraceWords * a.n_items, 1ull);
                return;
            }
        }
    }
    for (;;) {
        int k = 0;
        if (threadIdx.x == 0) {
            const unsigned long long c = atomicAdd(a.items, 1ull);
            k = (c < static_cast<unsigned long long>(a.n_items)) ? static_cast<int>(c) : -1;
        }
        k = __shfl(k, 0);
        if (k < 0) break;
        const uint64_t t_take = a.trace ? wall_ticks() : 0;
        lin_item<PH, Coh>(a, k / npb, (k % npb) * kLinItem + static_cast<int>(threadIdx.x), sln);
        if (a.trace) {
            __builtin_amdgcn_wave_barrier();
            if (threadIdx.x == 0) {
                unsigned long long* r = a.trace + static_cast<size_t>(kTraceWords) * k;
                r[0] = t_take;
                r[1] = wall_ticks();
                r[2] = static_cast<unsigned long long>(a.tag);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// adjoint window, phase 2: reverse-time transposed solves (B28, B27, B24)
//
// Reverse step s of reach p: gx = masked dL/d out[t+1]; gb = gx - sum over
// downstream d of adj * P_d; P_p = -c1 gb and G_p = c2 gb go to the upstream
// reaches; dL/d out[t] = direct + (gb * dgq/dgb + sum_d G_d * adj), masked by
// out[t] > lb when it is passed through the clamp of the previous step.

template <int RPL>
__global__ __launch_bounds__(64, 6) void k_adj_sweep(SweepArgs a) {
    using Wd = Width<RPL>;
    constexpr int CW = Wd::CW;
    __shared__ float ldsP[CW], ldsG[CW];
    __shared__ float stP[Wd::kStageDn], stG[Wd::kStageDn], sadj[Wd::kStageDn];
    __shared__ int spos[Wd::kStageDn];
    __shared__ short skind[Wd::kStageDn];
    // the step's records {-c1, c2, dgq/dgb, mask of dL/d out[t]} and direct
    // dL/d out[t], loaded straight into LDS
    __shared__ float4 prec[CW];
    __shared__ float pdq[CW];
    __shared__ int pgs[CW];  // gauge slot of the position, -1: none
    __shared__ float slam[CW];
    __shared__ unsigned char pvalid[CW];
    __shared__ short pe0[CW], pe1[CW], psub[CW];
    __shared__ short srep[kRepMax];
    __shared__ int lkc[kFastLinksDn][CW];
    __shared__ float lka[kFastLinksDn][CW];
    // the polled (P, G) pairs of the staged links, loaded straight into LDS
    __shared__ u32x4 spair[Wd::kStageDn];
    __shared__ int claimed;
    constexpr int kSlotsDn = Wd::kStageDn / 64;
    // vector memory operations a step issues after its polls: the pair
    // publish and the gb store of each row of the chunk
    constexpr int kStoresAfterPolls = 2 * RPL;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int P = sc.P;
    const int c = claim_chunk<-1>(a.claim, a.ch.C, &claimed, a.err);
    if (c < 0) return;
    const uint32_t epoch = launch_epoch(a.claim);
    bool unit = true;  // every diagonal of this step's I - c1 N is exactly 1
    const int rep0 = a.ch.rep_dn_ptr[c];
    const int nrep_all = a.ch.rep_dn_ptr[c + 1] - rep0;
    const int nrep = (nrep_all <= kRepMax) ? nrep_all : -1;
    const int p0 = c * CW;
    const int base = sc.dn_ptr[p0];
    const int nlinks = sc.dn_ptr[p0 + CW] - base;
    const int nsub = a.ch.nsubr[c];

    // A step's records and its cross-chunk inputs are loaded by issue(s),
    // which the previous step calls before its stores; all of them go from
    // memory straight to LDS. The stores that follow them are exactly
    // kStoresAfterPolls operations, so check() waits for the loads alone
    // (wait_vm) and not for the stores to reach the coherence point; a
    // register carrying a load across the step would make the compiler wait
    // for everything in flight. The tags of a pair name its row's parity
    // (row_tag), so a pair the previous step left in LDS never passes the check.
    bool stores_after_polls = false;
    // rows r of the chunk (positions r * 64 ...) holding a gauge: the others
    // read no gauge-sparse dL/dQ
    unsigned grows = 0;

    // link slots k this lane polls (bit k: staged link lane + 64 k is from
    // another chunk) and the positions it polls (0 where it has no link)
    unsigned pollm = 0;
    int dpos[kSlotsDn];
    int rpos = 0;  // the position this lane probes (lane < nrep)
    // per-lane operands of the solve, the same for every step: the position's
    // sub-level (-1: padding), its downstream links (CSR offset, count, first
    // kFastLinksDn codes and weights); and the gradient into x carried from
    // one step to the next (masked dL/d out[t+1])
    int my[RPL], e0[RPL], nl[RPL], code[RPL][kFastLinksDn];
    float w[RPL][kFastLinksDn], gx[RPL];
    bool tree = false;  // no position of the chunk has two downstream links
    // Step-invariant data, before the step loop (the per-step work of the
    // sweep is in issue and solve): the pointers used only here are dead
    // inside the loop.
    {
        stage_links<RPL>(c, base, base + nlinks, sc.dn_pos, sc.dn_adj, spos, sadj, skind);
        if (lane < nrep) srep[lane] = a.ch.rep_dn[rep0 + lane];
        if (a.gK) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int sl = a.gslot[p0 + r * 64 + lane];
                pgs[r * 64 + lane] = sl;
                if (__any(sl >= 0)) grows |= 1u << r;
            }
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const int p = p0 + j;
            pvalid[j] = sc.perm[p] >= 0;
            if (pvalid[j]) {
                const int ea = sc.dn_ptr[p], eb = sc.dn_ptr[p + 1];
                pe0[j] = static_cast<short>(ea - base);
                pe1[j] = static_cast<short>(eb - base);
                psub[j] = a.ch.subr[p];
                slam[j] = a.st_lam[p];
#pragma unroll
                for (int k = 0; k < kFastLinksDn; ++k) {
                    if (ea + k < eb) {
                        lkc[k][j] = link_code<RPL>(c, base, ea + k, sc.dn_pos[ea + k]);
                        lka[k][j] = sc.dn_adj[ea + k];
                    }
                }
            } else {
                psub[j] = -1;
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < kSlotsDn; ++k) {
            const int e = lane + 64 * k;
            const bool x = e < nlinks && skind[e] < 0;
            pollm |= x ? 1u << k : 0u;
            dpos[k] = x ? spos[e] : 0;
        }
        if (lane < nrep) rpos = spos[srep[lane]];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            my[r] = psub[j];
            e0[r] = (my[r] >= 0) ? pe0[j] : 0;
            nl[r] = (my[r] >= 0) ? pe1[j] - e0[r] : 0;
            gx[r] = (my[r] >= 0) ? slam[j] : 0.0f;
#pragma unroll
            for (int k = 0; k < kFastLinksDn; ++k) {
                code[r][k] = (k < nl[r]) ? lkc[k][j] : 0;
                w[r][k] = (k < nl[r]) ? lka[k][j] : 0.0f;
            }
        }
        bool one = true;
#pragma unroll
        for (int r = 0; r < RPL; ++r) one = one && nl[r] <= 1;
        tree = __all(one);
    }
    auto prepare = [&](int, bool) {};

    // Every lane loads, whether or not it needs the value (an empty position
    // reads its unused record slot, a lane without a cross-chunk link the
    // words of position 0): a load under a lane condition merges with the old
    // value afterwards, and that merge waits for the load at once.
    auto issue = [&](int s, bool again) {
        const size_t row = static_cast<size_t>(a.s_base + s) * P;
        if (!again) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const size_t p = static_cast<size_t>(p0 + r * 64 + lane);
                load_to_lds(a.rec + row + p, &prec[r * 64], 16);
                if (a.gqs && (!a.gK || ((grows >> r) & 1u))) {
                    // dense rows of P, or gauge rows of gK read at the position's slot
                    // (every lane of a row with a gauge loads: slot 0 stands in
                    // where there is no gauge)
                    const size_t idx = a.gK ? static_cast<size_t>(max(pgs[r * 64 + lane], 0)) : p;
                    load_to_lds(a.gqs + static_cast<size_t>(a.t0 + a.s_base + s) *
                                            static_cast<size_t>(a.gK ? a.gK : P) + idx,
                                &pdq[r * 64], 4);
                }
            }
        }
        // every lane of a used slot loads (position 0 where it has no link)
        const uint64_t* mrow = a.amsg + row * 2;
#pragma unroll
        for (int k = 0; k < kSlotsDn; ++k) {
            if (64 * k < nlinks) poll_pair_to_lds(mrow + 2 * dpos[k], &spair[64 * k]);
        }
        stores_after_polls = false;
    };
    auto check = [&](int s) -> bool {
        if (stores_after_polls) wait_vm<kStoresAfterPolls>();
        else wait_vm<0>();
        const uint32_t tag = row_tag(epoch, a.s_base + s);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kSlotsDn; ++k) {
            const int e = lane + 64 * k;
            if (64 * k < nlinks && ((pollm >> k) & 1u)) {
                const u32x4 w = lds_read_dma16(&spair[e]);
                if (!pair_ready(w, tag)) {
                    ok = false;
                } else {
                    stP[e] = __uint_as_float(w.x);
                    stG[e] = __uint_as_float(w.z);
                }
            }
        }
        const bool all = __all(ok);
        wave_lds_sync();
        return all;
    };

    // the G word of a link is published with its P word, in one access
    auto probe = [&](int s) -> bool {
        if (nrep <= 0) return true;  // no representatives: read the links again
        const uint64_t* row = a.amsg + static_cast<size_t>(a.s_base + s) * P * 2;
        const uint32_t tag = row_tag(epoch, a.s_base + s);
        bool ok = true;
        if (lane < nrep) ok = word_ready(poll(row + 2 * rpos + 1), tag);
        return __all(ok);
    };

    auto solve = [&](int s, bool last, auto&& before_stores) {
        const int sg = a.s_base + s;
        const auto mrow = pair_row(a.amsg + static_cast<size_t>(sg) * P * 2, P);
        const uint32_t tag = row_tag(epoch, sg);
        // the step's records (and gauge dL/dQ) were loaded into LDS before the
        // inputs check() waited for, so they have arrived (loads complete in
        // order); read once per step
        float4 rcv[RPL];
        float dqv[RPL];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            const u32x4 q = lds_read_dma16(&prec[j]);
            rcv[r] = make_float4(__uint_as_float(q.x), __uint_as_float(q.y), __uint_as_float(q.z),
                                 __uint_as_float(q.w));
            // direct dL/d out[t] of the position (zero where it has no gauge)
            const bool hasdq = a.gqs && (!a.gK || (((grows >> r) & 1u) && pgs[j] >= 0));
            dqv[r] = hasdq ? lds_read_dma4(&pdq[j]) : 0.0f;
        }
        {
            bool ok_diag = true;
#pragma unroll
            for (int r = 0; r < RPL; ++r)
                if (my[r] >= 0) ok_diag = ok_diag && (1.0f + rcv[r].x * 0.0f == 1.0f);
            unit = __all(ok_diag);
        }
        float gbv[RPL], lamv[RPL], gsum[RPL], pmv[RPL], gmv[RPL];
        // the staged (other chunks') values of the fast links, read before the
        // sub-levels: only links inside the chunk are read within them
        float xP[RPL][kFastLinksDn], xG[RPL][kFastLinksDn];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            gbv[r] = 0.0f;
            lamv[r] = 0.0f;
            gsum[r] = 0.0f;
            pmv[r] = 0.0f;
            gmv[r] = 0.0f;
#pragma unroll
            for (int k = 0; k < kFastLinksDn; ++k) {
                const int cd = code[r][k];
                xP[r][k] = (k < nl[r] && cd < 0) ? stP[-1 - cd] : 0.0f;
                xG[r][k] = (k < nl[r] && cd < 0) ? stG[-1 - cd] : 0.0f;
                // the reads are complete here: within the sub-levels no wait
                // for them holds back a sub-level's read behind the previous
                // sub-level's write
                asm volatile("" : "+v"(xP[r][k]), "+v"(xG[r][k]));
            }
        }
        if (tree) {
            // every position has at most one downstream link (a tree): the
            // sub-levels carry only P; G is gathered after them, below
            for (int l = 0; l < nsub; ++l) {
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    if (my[r] == l) {
                        const int j = r * 64 + lane;
                        const int cd = code[r][0];
                        // an in-bounds read for every lane (its own slot where
                        // its link leaves the chunk)
                        const float Pin = ldsP[(cd >= 0) ? cd : j];
                        const float Pd = (cd >= 0) ? Pin : xP[r][0];
                        float sumP = 0.0f;
                        if (nl[r] > 0) sumP = sumP + w[r][0] * Pd;       // B27
                        const float gb = unit ? (gx[r] - sumP)
                                              : div_diag(gx[r] - sumP, 1.0f + rcv[r].x * 0.0f);
                        const float Pm = rcv[r].x * gb;
                        ldsP[j] = Pm;
                        pmv[r] = Pm;
                        gmv[r] = rcv[r].y * gb;
                        gbv[r] = gb;
                    }
                }
                wave_lds_sync();
            }
#pragma unroll
            for (int r = 0; r < RPL; ++r) ldsG[r * 64 + lane] = gmv[r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (my[r] >= 0 && nl[r] > 0) {
                    const int cd = code[r][0];
                    const float Gd = (cd >= 0) ? ldsG[cd] : xG[r][0];
                    gsum[r] = 0.0f + Gd * w[r][0];                      // B24 N^T gi
                }
            }
        } else for (int l = 0; l < nsub; ++l) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (my[r] == l) {
                    const int j = r * 64 + lane;
                    const float4 rc = rcv[r];  // {-c1, c2, dgq/dgb, mask}
                    float sumP = 0.0f, sumG = 0.0f;
                    if (nl[r] <= kFastLinksDn) {
#pragma unroll
                        for (int k = 0; k < kFastLinksDn; ++k) {
                            if (k < nl[r]) {
                                const int cd = code[r][k];
                                const float Pd = (cd >= 0) ? ldsP[cd] : xP[r][k];
                                const float Gd = (cd >= 0) ? ldsG[cd] : xG[r][k];
                                sumP = sumP + w[r][k] * Pd;              // B27
                                sumG = sumG + Gd * w[r][k];              // B24 N^T gi
                            }
                        }
                    } else {
                        for (int e = e0[r]; e < e0[r] + nl[r]; ++e) {
                            const int kd = skind[e];
                            const float Pd = (kd >= 0) ? ldsP[kd] : stP[e];
                            const float Gd = (kd >= 0) ? ldsG[kd] : stG[e];
                            const float adj = sadj[e];
                            sumP = sumP + adj * Pd;
                            sumG = sumG + Gd * adj;
                        }
                    }
                    const float gb = unit ? (gx[r] - sumP)
                                          : div_diag(gx[r] - sumP, 1.0f + rc.x * 0.0f);
                    const float Pm = rc.x * gb;
                    const float Gm = rc.y * gb;
                    ldsP[j] = Pm;
                    ldsG[j] = Gm;
                    pmv[r] = Pm;
                    gmv[r] = Gm;
                    gbv[r] = gb;
                    gsum[r] = sumG;
                }
            }
            wave_lds_sync();
        }
        before_stores();
        // Published after the solve, as in the forward. Exactly
        // kStoresAfterPolls vector memory operations follow the polls: every
        // lane stores (padding positions' words are never read).
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int j = r * 64 + lane;
            publish_pair(mrow, p0 + j, pmv[r], gmv[r], tag);
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r)
            a.gbuf[static_cast<size_t>(sg) * P + p0 + r * 64 + lane] = gbv[r];
        stores_after_polls = true;
        // B24 for every position of the chunk at once, after the publish
#pragma unroll
        for (int r = 0; r < RPL; ++r)
            if (my[r] >= 0)
                lamv[r] = (rcv[r].w != 0.0f) ? dqv[r] + (gbv[r] * rcv[r].z + gsum[r]) : 0.0f;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            if (my[r] >= 0) {
                gx[r] = lamv[r];
                if (last) a.st_lam[p0 + r * 64 + lane] = lamv[r];
            }
        }
    };

    chunk_loop<-1>(a.nsteps, a.err, a.trace.rec ? a.trace.rec + kTraceWords * c : nullptr, issue,
                   prepare, check, probe, solve);
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
        const float4 j = make_float4(a.recj[o], a.recj[a.plane + o], a.recj[2 * a.plane + o],
                                     a.gqp_out ? a.recj[3 * a.plane + o] : 0.0f);
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
    const int c = claim_chunk<+1>(a.claim, a.ch.C, &claimed, a.err);
    if (c < 0) return;
    const uint32_t epoch = launch_epoch(a.claim);
    const int base = sc.up_ptr[c * 64];
    const int end = sc.up_ptr[c * 64 + 64];
    auto prepare = [&](int, bool) {};
    auto ready = [&](int) -> bool {
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int u = sc.up_pos[e];
            if ((u >> 6) != c) {
                const uint64_t wv = poll(a.slot + u);
                if (!word_ready(wv, epoch)) ok = false;
                else stage[e - base] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };
    auto solve = [&](int, bool, auto&&) {
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
                publish(a.slot + p, x, epoch);
            }
            __syncthreads();
        }
        if (valid) a.x[sc.perm[p]] = a.clamp ? fmaxr(x, a.lb) : x;
    };
    auto issue = [&](int, bool) {};
    auto probe = [&](int) { return true; };
    chunk_loop<+1>(1, a.err, nullptr, issue, prepare, ready, probe, solve);
}

__global__ __launch_bounds__(64) void k_tri_adj(TriArgs a) {
    __shared__ float lds_g[64];
    __shared__ float stage[kStageDnBase];
    __shared__ int claimed;
    const int lane = threadIdx.x;
    const Sched& sc = a.sc;
    const int c = claim_chunk<-1>(a.claim, a.ch.C, &claimed, a.err);
    if (c < 0) return;
    const uint32_t epoch = launch_epoch(a.claim);
    const int base = sc.dn_ptr[c * 64];
    const int end = sc.dn_ptr[c * 64 + 64];
    auto prepare = [&](int, bool) {};
    auto ready = [&](int) -> bool {
        bool ok = true;
        for (int e = base + lane; e < end; e += 64) {
            const int d = sc.dn_pos[e];
            if ((d >> 6) != c) {
                const uint64_t wv = poll(a.slot + d);
                if (!word_ready(wv, epoch)) ok = false;
                else stage[e - base] = word_value(wv);
            }
        }
        const bool all = __all(ok);
        __syncthreads();
        return all;
    };
    auto solve = [&](int, bool, auto&&) {
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
                publish(a.slot + p, g, epoch);
            }
            __syncthreads();
        }
        if (valid) a.x[sc.perm[p]] = g;
    };
    auto issue = [&](int, bool) {};
    auto probe = [&](int) { return true; };
    chunk_loop<-1>(1, a.err, nullptr, issue, prepare, ready, probe, solve);
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

// Rows [t_first, t_first + rows) of q ([..][R], reach order) into schedule
// order (out [rows][P], 0 at padding positions). Workgroups are dealt to the
// 8 XCDs in turn (workgroup b on XCD b % 8), and each XCD takes whole rows
// (row k on XCD k % 8, its workgroups one after another): the gathers of a
// row then hit the one L2 that holds its lines (1.4 MB for 350k reaches)
// instead of every XCD's L2 fetching every row.
constexpr int kXcds = 8;
__global__ __launch_bounds__(256) void k_permute_rows(Sched sc, const float* q, int t_first,
                                                      int rows, int row_blocks, float* out) {
    const int x = static_cast<int>(blockIdx.x % kXcds);
    const int local = static_cast<int>(blockIdx.x / kXcds);
    const int k = (local / row_blocks) * kXcds + x;
    const int p = (local % row_blocks) * 256 + static_cast<int>(threadIdx.x);
    if (k >= rows || p >= sc.P) return;
    const int o = sc.perm[p];
    out[static_cast<size_t>(k) * sc.P + p] =
        (o >= 0) ? q[static_cast<size_t>(t_first + k) * sc.R + o] : 0.0f;
}

__global__ void k_gauge_row0(const int* reach, int K, const float* q0, float lb, float* Qg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) Qg[i] = fmaxr(q0[reach[i]], lb);
}
__global__ void k_gauge_scatter_row(const int* reach, int K, const float* g, float* dense) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) dense[reach[i]] = g[i];
}
__global__ void k_zero_f(size_t n, float* x) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}
__global__ void k_zero_u64(size_t n, unsigned long long* x) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0ull;
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

void launch_begin(LaunchState* ls, const unsigned long long* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_launch, dim3(1), dim3(64), 0, s, ls, ctr);
}
void launch_fwd(const FwdArgs& a, int variant, int rpl, int grid, hipStream_t s) {
    if (rpl == 2) launch_fwd_r<2>(a, variant, grid, s);
    else launch_fwd_r<1>(a, variant, grid, s);
}

static inline dim3 g1(size_t n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

int lin_items(int nsteps, int P) {
    return ((nsteps + kLinSteps - 1) / kLinSteps) * ((P + kLinItem - 1) / kLinItem);
}

void launch_lin(const LinArgs& a, int variant, int grid, hipStream_t s) {
    if (!a.items) {
        const dim3 g(static_cast<unsigned>((a.sc.P + kLinBlock - 1) / kLinBlock),
                     static_cast<unsigned>((a.nsteps + kLinSteps - 1) / kLinSteps));
        const dim3 b(kLinBlock);
        switch (variant) {
            case PH_DDR: hipLaunchKernelGGL(k_adj_lin<PH_DDR>, g, b, 0, s, a); break;
            case PH_POS: hipLaunchKernelGGL(k_adj_lin<PH_POS>, g, b, 0, s, a); break;
            default: hipLaunchKernelGGL(k_adj_lin<PH_DEFAULT>, g, b, 0, s, a); break;
        }
        return;
    }
    const dim3 g(static_cast<unsigned>(grid)), b(kLinItem);
    const bool coh = a.fwd_epoch != 0;
    switch (variant) {
        case PH_DDR:
            if (coh) hipLaunchKernelGGL((k_adj_lin_items<PH_DDR, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_adj_lin_items<PH_DDR, false>), g, b, 0, s, a);
            break;
        case PH_POS:
            if (coh) hipLaunchKernelGGL((k_adj_lin_items<PH_POS, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_adj_lin_items<PH_POS, false>), g, b, 0, s, a);
            break;
        default:
            if (coh) hipLaunchKernelGGL((k_adj_lin_items<PH_DEFAULT, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_adj_lin_items<PH_DEFAULT, false>), g, b, 0, s, a);
            break;
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
void launch_permute_rows(const Sched& sc, const float* q, int t_first, int rows, float* out,
                         hipStream_t s) {
    const int row_blocks = (sc.P + 255) / 256;
    // A launch's grid is at most 2^32 - 1 work-items (the dispatch packet
    // holds the count in 32 bits; a larger grid is cut short without an
    // error): launches of at most 2^24 - 1 workgroups, in whole turns of
    // kXcds rows.
    const size_t max_blocks = (size_t{1} << 24) - 1;
    const int per = static_cast<int>(
        std::max<size_t>(1, max_blocks / (static_cast<size_t>(row_blocks) * kXcds)) * kXcds);
    for (int k0 = 0; k0 < rows; k0 += per) {
        const int n = std::min(per, rows - k0);
        const size_t blocks = static_cast<size_t>((n + kXcds - 1) / kXcds) * kXcds * row_blocks;
        hipLaunchKernelGGL(k_permute_rows, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, sc,
                           q, t_first + k0, n, row_blocks, out + static_cast<size_t>(k0) * sc.P);
    }
}
void launch_zero(size_t n, float* x, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_zero_f, g1(n), dim3(256), 0, s, n, x);
}
void launch_gauge_row0(const int* reach, int K, const float* q0, float lb, float* Qg,
                       hipStream_t s) {
    if (K > 0) hipLaunchKernelGGL(k_gauge_row0, g1(K), dim3(256), 0, s, reach, K, q0, lb, Qg);
}
void launch_gauge_scatter_row(const int* reach, int K, const float* g, float* dense,
                              hipStream_t s) {
    if (K > 0) hipLaunchKernelGGL(k_gauge_scatter_row, g1(K), dim3(256), 0, s, reach, K, g, dense);
}
void launch_zero_u64(size_t n, unsigned long long* x, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_zero_u64, g1(n), dim3(256), 0, s, n, x);
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

// Internal types shared by the kernels (ddrs_kernels.hip) and the host side
// (ddrs_host.hip). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include "../../include/ddrs_hip.h"
#include "mc_physics.h"

namespace ddrs {

// Links of a chunk staged in LDS, per 64 reaches of chunk width: upstream
// links (forward) and downstream links (adjoint; one per reach in a tree, so
// a tree's chunks never close early for them). The kernels' LDS is sized so
// that 24 one-wave workgroups (six per SIMD at 80 VGPRs) fit
// in a CU's 160 KB: persistent launches need every chunk of the window
// resident at once for the time skew to pipeline.
constexpr int kStageUpBase = 192;
constexpr int kStageDnBase = 64;
// Chunk widths: 64 * RPL positions for RPL (reaches per lane) in {1, 2}.
constexpr int kNumWidths = 2;
// Longest chain of links inside a 64-position chunk (its sub-levels): every
// time step walks it serially, so it bounds the step rate of the chunk.
constexpr int kMaxSubLevels = 8;
// Positions per block of the schedule layout (64 or 128): the sub-level cap
// holds per block, and so for chunks of the block's width or narrower.
constexpr int kLayoutBlock = 64;

// Level-major schedule, device side. Position p in [0, P); reach-indexed
// (reference order) arrays have R entries.
struct Sched {
    int R, P;
    const int* perm;          // [P] position -> reach id, -1 for an empty position
    const int* up_ptr;        // [P+1] upstream entries of position p
    const int* up_pos;        // [E] upstream position, reference CSR order
    const float* up_adj;      // [E] N value
    const int* up_k;          // [E] CsrPattern nonzero index
    const int* dn_ptr;        // [P+1] downstream entries of position p
    const int* dn_pos;        // [E] downstream position, transposed-CSR order
    const float* dn_adj;      // [E]
    const int* dn_k;          // [E] CsrPattern nonzero index
    const int* diag_k;        // [P] CsrPattern index of the diagonal
};

// The schedule cut into chunks of 64 << (shift - 6) consecutive positions; a
// chunk is one workgroup of one wavefront.
struct Chunks {
    int C, shift;
    const unsigned char* sub;   // [P] forward sub-level inside the chunk
    const unsigned char* subr;  // [P] reverse sub-level inside the chunk
    const int* nsub;            // [C]
    const int* nsubr;           // [C]
    // One representative link per producing chunk (link index relative to the
    // chunk's first link): polled while a chunk waits, instead of every link.
    const int* rep_up_ptr;      // [C+1]
    const short* rep_up;
    const int* rep_dn_ptr;      // [C+1]
    const short* rep_dn;
};

// Representatives a chunk keeps in LDS; a chunk fed by more chunks polls all
// of its links while it waits.
constexpr int kRepMax = 64;

// Per-launch values of a persistent launch, written on the device by a
// one-thread kernel enqueued right before it, so that a launch captured in a
// graph replays with fresh values: the claim counter at the start of the
// launch, and the epoch that tags the words the launch publishes.
struct LaunchState {
    unsigned long long claim_base;
    uint32_t epoch;
    uint32_t pad;
};

// Chunk claims of a persistent launch: the workgroup that increments *ctr to
// ls->claim_base + k takes the k-th chunk in dependency order.
struct Claim {
    unsigned long long* ctr;
    const LaunchState* ls;
};

// Optional per-chunk trace of a persistent launch (null: off): for chunk c,
// rec[8c..8c+7] = wall clock (100 MHz ticks) when the workgroup claimed c,
// when its first step was solved, when its last step was solved, the number
// of polls that found an input missing, and the ticks spent in prepare, in
// the first poll of each step, waiting for late inputs, and in solve.
constexpr int kTraceWords = 8;
struct Trace {
    unsigned long long* rec;
};

struct FwdArgs {
    Sched sc;
    Chunks ch;
    Claim claim;
    Trace trace;
    const float4* st;      // [P][3] reach statics, schedule order
    Phys P;
    int t0;          // out[t0] is the window's initial state
    int s_base;      // window-local index of this launch's first step
    int nsteps;      // steps in this launch
    const float* qps;      // [T-1][P] q' in schedule order
    float* Q_out;          // [T][R] reference order or null
    float* X_out;          // [R] pre-clamp solve (single step) or null
    // Gauge readout (null: off): Qg[t][slot] for the positions whose gslot
    // (schedule order, -1: none) is a slot in [0, gK).
    float* Qg;
    const int* gslot;
    int gK;
    uint64_t* hist;        // [window][P] tagged x, schedule order
    // [window][P] upstream sums of each reach-step, or null: (sum_k adj_k x[up k],
    // sum_k adj_k max(x[up k], lb)) over the reach's upstream links in CSR order
    float2* ihist;
    float* st_q;           // [P] out[t] of the current step
    float* st_i;           // [P] N out[t]
    int* err;
    // chunks at or beyond this index run their coefficients and input checks
    // at a raised issue priority (results do not depend on it)
    int prio_from;
};

// Adjoint, phase 1 (every reach-step of a segment independently): the step's
// coefficients and its backward for gb = 1, from the forward history.
struct LinArgs {
    Sched sc;
    const float4* st;
    Phys P;
    int t0, s_base, nsteps;
    const float* qps;      // [T-1][P] q' in schedule order
    int mask_qp;           // 1: gradient w.r.t. raw q' through max(q', lb)
    int mask_lam;          // 1: dL/d out[t] is passed on through out[t] = max(x, lb)
    const uint64_t* hist;  // forward history of the segment
    // the forward's upstream sums of the segment (FwdArgs::ihist), or null:
    // then they are gathered from the upstream positions' history words
    const float2* ihist;
    const float* ck_q;     // [P] out[t0] (schedule order)
    const float* ck_i;     // [P] N out[t0]
    float4* rec;           // [window][P] {-c1, c2, d gq / d gb, mask of dL/d out[t]}
    // d(gn, gqsp, gp, gq') / d gb of each reach-step, in four planes of
    // `plane` floats ([window][P] each): the finishing phase reads the fourth
    // only when dL/dq' is asked for
    float* recj;
    size_t plane;
    // With items == null the grid holds one workgroup per 256 positions and 32
    // steps. Otherwise one-wave workgroups take items (64 positions, 32 steps)
    // in step-block-major order from the counter *items (zero before the first
    // launch that uses it) until n_items are taken.
    unsigned long long* items;
    int n_items;
    // Linearising while the forward window still runs (fwd_epoch != 0): the
    // history is read with device-coherent loads and every word checked for
    // the forward's epoch (awaited when missing). A workgroup starts taking
    // items only once every forward workgroup has claimed its chunk
    // (*fwd_claim_ctr >= fwd_claims), and leaves its share to a later launch
    // on the same counter if the claims stop advancing for a while.
    uint32_t fwd_epoch;
    const unsigned long long* fwd_claim_ctr;
    unsigned long long fwd_claims;
    int* err;
    // Optional trace of a launch on the counter (null: off): for item k,
    // trace[8k..8k+2] = wall clock when taken and when done, and `tag`; the
    // record after the last item counts workgroups that left without an item.
    unsigned long long* trace;
    int tag;
};

// Adjoint, phase 2 (persistent, reverse time): the transposed solves.
struct SweepArgs {
    Sched sc;
    Chunks ch;
    Claim claim;
    Trace trace;
    int t0, s_base, nsteps;
    // dL/dQ (direct), or null: with gK == 0 [T-1][P] in schedule order; with
    // gK > 0 gauge-sparse [T-1][gK], read at the slot gslot[p] of the position
    // (zero where gslot is -1)
    const float* gqs;
    const int* gslot;
    int gK;
    const float4* rec;
    uint64_t* amsg;        // [window][P][2] tagged (P, G)
    float* st_lam;         // [P] gradient into x of the step (masked dL/d out[t+1])
    float* gbuf;           // [window][P] gb of every reach-step
    int* err;
};

// Adjoint, phase 3: parameter and q' gradients from gb.
struct FinArgs {
    int R, P;
    const int* perm;
    int t0, s_base, nsteps;
    const float* recj;     // four planes (LinArgs::recj)
    size_t plane;
    const float* gbuf;
    float* acc_n;
    float* acc_qsp;
    float* acc_p;
    float* gqp_out;        // [T][R] reference order or null
};

struct TriArgs {
    Sched sc;
    Chunks ch;             // 64-wide
    Claim claim;
    const float* a_values;  // [nnz] or null (A = I - N)
    const float* b;         // [R] reference order
    const float* mask_x;    // [R] or null: gradient masked where mask_x <= lb
    float* x;               // [R] reference order
    uint64_t* slot;         // [P]
    int clamp;
    float lb;
    int* err;
};

// The LaunchState of the next persistent launch: claim base from *ctr, epoch + 1.
void launch_begin(LaunchState* ls, const unsigned long long* ctr, hipStream_t s);
// Persistent launches run one workgroup per chunk (grid = Chunks::C).
void launch_fwd(const FwdArgs& a, int variant, int rpl, int grid, hipStream_t s);
// grid: one workgroup per item (a.items null), else `grid` workgroups
void launch_lin(const LinArgs& a, int variant, int grid, hipStream_t s);
// linearisation work items of a launch of nsteps steps over P positions
int lin_items(int nsteps, int P);
void launch_zero_u64(size_t n, unsigned long long* x, hipStream_t s);
void launch_sweep(const SweepArgs& a, int rpl, int grid, hipStream_t s);
void launch_fin(const FinArgs& a, hipStream_t s);
void launch_tri_fwd(const TriArgs& a, int grid, hipStream_t s);
void launch_tri_adj(const TriArgs& a, int grid, hipStream_t s);
void launch_grad_a(int nnz, const int* row, const int* col, const float* gb, const float* x,
                   float* ga, hipStream_t s);
void launch_bind_statics(const Sched& sc, const float* n, const float* qsp, const float* psp,
                         const float* len, const float* slope, const float* xst, float slope_lb,
                         float4* st, hipStream_t s);
void launch_state(const Sched& sc, const float* q, int clamp, float lb, float* st_q,
                  float* st_i, float* Q_row0, hipStream_t s);
void launch_hist_from_x(const Sched& sc, const float* x, uint64_t* hist, hipStream_t s);
void launch_lam_init(const Sched& sc, const float* g, const uint64_t* hist_row, float lb,
                     float* lam, float* a0, float* a1, float* a2, hipStream_t s);
void launch_unpermute(const Sched& sc, const float* src, float* dst, hipStream_t s);
void launch_grad_q0(const Sched& sc, const float* lam, const float* q0, float lb, float* g,
                    hipStream_t s);
void launch_zero(size_t n, float* x, hipStream_t s);
// Qg[s] = max(q0[reach[s]], lb) for s < K (row 0 of a gauge readout)
void launch_gauge_row0(const int* reach, int K, const float* q0, float lb, float* Qg,
                       hipStream_t s);
// dense[reach[s]] = g[s] for s < K (other entries untouched)
void launch_gauge_scatter_row(const int* reach, int K, const float* g, float* dense,
                              hipStream_t s);
// out[k][p] = q[(t_first + k) * R + perm[p]] for k < rows (0 at empty positions)
void launch_permute_rows(const Sched& sc, const float* q, int t_first, int rows, float* out,
                         hipStream_t s);

// Resident workgroups per CU of a persistent kernel.
enum OccKernel : int { OCC_FWD = 0, OCC_SWEEP = 1, OCC_TRI_FWD = 2, OCC_TRI_ADJ = 3 };
int occupancy_blocks(int kernel, int variant, int rpl);

}  // namespace ddrs

// Internal types shared by the kernels (ddrs_kernels.hip) and the host side
// (ddrs_host.hip). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include "../../include/ddrs_hip.h"
#include "mc_physics.h"

namespace ddrs {

// Upstream (downstream) entries of one 64-reach chunk are staged in LDS.
constexpr int kStageMax = 256;
// Chunks one persistent workgroup may own (C <= kMaxChunksPerWave * grid).
constexpr int kMaxChunksPerWave = 128;

// Level-major schedule, device side. Position p in [0, R).
struct Sched {
    int R, C;
    const int* perm;          // [R] position -> reach id
    const int* up_ptr;        // [R+1] upstream entries of position p
    const int* up_pos;        // [E] upstream position, reference CSR order
    const float* up_adj;      // [E] N value
    const int* up_k;          // [E] CsrPattern nonzero index
    const int* dn_ptr;        // [R+1] downstream entries of position p
    const int* dn_pos;        // [E] downstream position, transposed-CSR order
    const float* dn_adj;      // [E]
    const int* dn_k;          // [E] CsrPattern nonzero index
    const int* diag_k;        // [R] CsrPattern index of the diagonal
    const unsigned char* sub;   // [R] forward sub-level inside the chunk
    const unsigned char* subr;  // [R] reverse sub-level inside the chunk
    const int* nsub;          // [C]
    const int* nsubr;         // [C]
};

// Per-reach statics in schedule order.
struct Statics {
    float* n;
    float* qsp;
    float* p;
    float* len;
    float* slope;   // clamped
    float* sqrt_s;
    float* xst;
};

struct FwdArgs {
    Sched sc;
    Statics st;
    Phys P;
    int t0;          // out[t0] is the window's initial state
    int s_base;      // window-local index of this launch's first step
    int nsteps;      // steps in this launch
    uint32_t epoch;
    const float* qprime;   // [T][R] reference order
    float* Q_out;          // [T][R] reference order or null
    float* X_out;          // [R] pre-clamp solve (single step) or null
    uint64_t* hist;        // [window][R] tagged x, schedule order
    float* st_q;           // [R] out[t] of the current step
    float* st_i;           // [R] N out[t]
    int* err;
};

struct AdjArgs {
    Sched sc;
    Statics st;
    Phys P;
    int t0, s_base, nsteps;
    uint32_t epoch;
    const float* qprime;   // [T][R]
    const float* gQ;       // [T][R] or null
    float* gqp_out;        // [T][R] or null
    int mask_qp;           // 1: gradient w.r.t. raw q' through max(q', lb)
    const uint64_t* hist;  // forward history of the window
    const float* ck_q;     // out[t0] (schedule order)
    const float* ck_i;     // N out[t0]
    uint64_t* amsg;        // [window][R][2] tagged (P, G)
    float* st_lam;         // [R] dL/d out[t+1] on entry of step t
    float* acc_n;
    float* acc_qsp;
    float* acc_p;
    float* sc_q;           // [R] out[t] of the step being processed (prepare -> solve)
    float* sc_i;           // [R] N out[t]
    float* sc_qp;          // [R] raw q' of the step
    int* err;
};

struct TriArgs {
    Sched sc;
    uint32_t epoch;
    const float* a_values;  // [nnz] or null (A = I - N)
    const float* b;         // [R] reference order
    const float* mask_x;    // [R] or null: gradient masked where mask_x <= lb
    float* x;               // [R] reference order
    uint64_t* slot;         // [R]
    int clamp;
    float lb;
    int* err;
};

void launch_fwd(const FwdArgs& a, const ddrs_params* p, int grid, hipStream_t s);
void launch_adj(const AdjArgs& a, const ddrs_params* p, int grid, hipStream_t s);
void launch_tri_fwd(const TriArgs& a, int grid, hipStream_t s);
void launch_tri_adj(const TriArgs& a, int grid, hipStream_t s);
void launch_grad_a(int nnz, const int* row, const int* col, const float* gb, const float* x,
                   float* ga, hipStream_t s);
void launch_bind_statics(int R, const int* perm, const float* n, const float* qsp,
                         const float* psp, const float* len, const float* slope,
                         const float* xst, float slope_lb, const Statics& st, hipStream_t s);
void launch_state(const Sched& sc, const float* q, int clamp, float lb, float* st_q,
                  float* st_i, float* Q_row0, hipStream_t s);
void launch_hist_from_x(int R, const int* perm, const float* x, uint64_t* hist, hipStream_t s);
void launch_lam_init(int R, const int* perm, const float* g, float* lam, float* a0, float* a1,
                     float* a2, hipStream_t s);
void launch_unpermute(int R, const int* perm, const float* src, float* dst, hipStream_t s);
void launch_grad_q0(int R, const int* perm, const float* lam, const float* q0, float lb,
                    float* g, hipStream_t s);
void launch_zero(size_t n, float* x, hipStream_t s);
// which: 0..2 forward (default, ddr, pos), 3..5 adjoint, 6 tri fwd, 7 tri adj
int occupancy_blocks(int which);

}  // namespace ddrs

/*
 * ddrs_hip.h — C ABI of the MI355X (gfx950) Muskingum-Cunge routing library.
 *
 * Drop-in boundary for the reference ddrs path (taddyb/ddrs):
 *   - the BURN custom op TimestepOp : Backward<I,5> (src/routing/mmc_op.rs:850-893,
 *     timestep_forward 1704-1857), parents [n, q_spatial, p_spatial, q_t, q'_t];
 *   - the src/sparse seam it calls (src/sparse/dispatch.rs:57-244;
 *     triangular_csr_solve src/sparse/mod.rs:483-531, CsrSolveOp 425-472);
 *   - the time loop of MuskingumCunge::forward (src/routing/mmc.rs:519-632) and the
 *     reverse walk its autograd tape performs, as one window forward + adjoint.
 *
 * Conventions
 *   - Plain pointers and sizes only. Pointers named d_* are device pointers owned
 *     by the caller; pointers named h_* are host pointers.
 *   - Per-reach vectors are in the caller's (reference) reach order, length R.
 *     Time series are time-major [T][R] (row t holds all reaches at step t).
 *   - Every call that launches work takes a hipStream_t (passed as void*; NULL is
 *     the default stream) and returns ddrs_status; on failure ddrs_last_error()
 *     describes it (thread-local).
 *   - Nothing is allocated inside a step: workspaces are created up front.
 *   - Thread-compatible: one stream per workspace at a time.
 *   - All arithmetic is fp32, indices int32.
 */
#ifndef DDRS_HIP_H
#define DDRS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DDRS_OK = 0,
    DDRS_ERR_INVALID = 1,   /* bad argument / malformed adjacency */
    DDRS_ERR_HIP = 2,       /* HIP runtime failure */
    DDRS_ERR_CAPACITY = 3,  /* a size limit of this build was exceeded */
    DDRS_ERR_NOMEM = 4
} ddrs_status;

typedef struct ddrs_plan* ddrs_plan_t;
typedef struct ddrs_workspace* ddrs_workspace_t;

/* Scalar routing parameters (reference: config.rs AttributeMinimums 426-446,
 * Params 580-672, mmc.rs DT_SECONDS 33). */
typedef struct {
    float dt;               /* 3600 */
    float discharge_lb;     /* 1e-4 */
    float depth_lb;         /* 0.01 */
    float velocity_lb;      /* 0.01 */
    float bottom_width_lb;  /* 0.01 */
    float slope_lb;         /* 1e-3, applied by ddrs_mc_bind_statics (mmc.rs:259-264) */
    int32_t ddr_match;          /* 0 (default physics) or 1 (legacy 5/3 celerity, X = x_storage) */
    int32_t enforce_positivity; /* S18'/S19' clamp; ignored when ddr_match */
} ddrs_params;

/* Execution form of a window (multi-step) call. */
typedef enum {
    DDRS_FORM_PLAIN = 0,      /* one kernel launch per timestep */
    DDRS_FORM_PERSISTENT = 1  /* one persistent, time-skewed launch per window */
} ddrs_form;

const char* ddrs_last_error(void);
const char* ddrs_version(void);

/* ---------------------------------------------------------------------------
 * Plan: the network. Replaces CsrPattern::from_sparse + build_transposed_pattern
 * (sparse/mod.rs:173-304) and the cuSPARSE pattern cache
 * (sparse/cusparse.rs:1141-1850). COO with rows = downstream, cols = upstream,
 * rows > cols (SparseAdjacency, sparse/mod.rs:44-105), host arrays. Builds the
 * bit-exact CSR + transposed view on the host and a level-major device schedule.
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_plan_create(int32_t n, int32_t nnz_off, const int32_t* h_rows,
                             const int32_t* h_cols, const float* h_vals,
                             ddrs_plan_t* out);
/* As ddrs_plan_create, with the layout of the device schedule: blocks of
 * `block` positions (64 or 128; 0: 64) holding at most `max_sub_levels`
 * sub-levels (longest chain of links inside the block; 0: 8). Smaller caps
 * spread thin chains over more chunks, which pipeline over time steps.
 * Results do not depend on either. No reference counterpart. */
ddrs_status ddrs_plan_create_ex(int32_t n, int32_t nnz_off, const int32_t* h_rows,
                                const int32_t* h_cols, const float* h_vals,
                                int32_t max_sub_levels, int32_t block, ddrs_plan_t* out);
void ddrs_plan_destroy(ddrs_plan_t plan);
/* n, nnz (= nnz_off + n), number of 64-reach chunks, number of levels. */
ddrs_status ddrs_plan_info(ddrs_plan_t plan, int32_t* n, int32_t* nnz, int32_t* n_chunks,
                           int32_t* n_levels);
/* Copy the CsrPattern arrays (sparse/mod.rs:115-146) to host buffers sized
 * n+1 / nnz. Any pointer may be NULL. */
ddrs_status ddrs_plan_get_csr(ddrs_plan_t plan, int32_t* h_crow, int32_t* h_col,
                              int32_t* h_row_for_nnz, float* h_adj_values,
                              float* h_diag_mask, int32_t* h_trans_crow,
                              int32_t* h_trans_col, int32_t* h_trans_to_orig);
/* Copy the schedule: h_perm[p] = reach at schedule position p, h_level[i] = level
 * (longest upstream path) of reach i. */
ddrs_status ddrs_plan_get_schedule(ddrs_plan_t plan, int32_t* h_perm, int32_t* h_level);
/* Number of schedule positions P (>= n: chunks are padded to whole sub-levels). */
ddrs_status ddrs_plan_positions(ddrs_plan_t plan, int32_t* P);
/* q' in schedule order for DDRS_WS_SCHEDULE_FORCING workspaces: d_q_sched[k][p]
 * = d_q_prime[k][reach at position p] for k < rows (0 at padding positions);
 * d_q_prime [rows][n] in reach order, d_q_sched [rows][P], device pointers,
 * enqueued on `stream`. Called once per batch by the data side, so that no
 * training or routing step permutes its forcing. No reference counterpart
 * (the reference's forcing stays in reach order: mmc.rs:519-632 reads
 * q_prime[t-1] per step). */
ddrs_status ddrs_plan_permute_forcing(ddrs_plan_t plan, int32_t rows, const float* d_q_prime,
                                      float* d_q_sched, void* stream);

/* ---------------------------------------------------------------------------
 * Workspace: device state for windows of up to `window` timesteps (the
 * activation-checkpoint interval k: forward history, adjoint messages and
 * checkpoints are sized for it). max_steps bounds the total steps T-1 of one
 * window call (checkpoint storage). 288 GB MI355X: C2 (350k x 720) with
 * window 719 and the dense dL/dQ buffer uses 17.7 GiB.
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_workspace_create(ddrs_plan_t plan, int32_t window, int32_t max_steps,
                                  ddrs_workspace_t* out);
/* As ddrs_workspace_create with flags: DDRS_WS_NO_DENSE_GRAD leaves out the
 * [max_steps][n] buffer of a dense dL/dQ (only the _gauged adjoint is then
 * available; the dense one returns DDRS_ERR_INVALID). For windows whose q'
 * alone fills a large part of HBM. DDRS_WS_SCHEDULE_FORCING: the window calls
 * (ddrs_mc_window_fwd/_adj and their _gauged forms) take q' as [T][P] rows in
 * schedule order, written by ddrs_plan_permute_forcing, and read it in place:
 * the workspace holds no [max_steps][P] copy and the calls permute nothing.
 * The single-step calls keep reach order. Results are bit-identical to reach
 * order. No reference counterpart (the reference reads q' [T][n] in reach
 * order, mmc.rs:519-632). */
enum { DDRS_WS_NO_DENSE_GRAD = 1, DDRS_WS_SCHEDULE_FORCING = 2 };
ddrs_status ddrs_workspace_create_ex(ddrs_plan_t plan, int32_t window, int32_t max_steps,
                                     int32_t flags, ddrs_workspace_t* out);
void ddrs_workspace_destroy(ddrs_workspace_t ws);
/* Bytes of device memory the workspace holds. */
ddrs_status ddrs_workspace_bytes(ddrs_workspace_t ws, size_t* bytes);
/* Bytes ddrs_workspace_create_ex(plan, window, max_steps, flags) would
 * allocate (before the gauge buffers of ddrs_workspace_set_gauges), without
 * allocating: for choosing a checkpoint window that fits the memory at hand.
 * No reference counterpart (the reference keeps every step's state). */
ddrs_status ddrs_workspace_query_bytes(ddrs_plan_t plan, int32_t window, int32_t max_steps,
                                       int32_t flags, size_t* bytes);
/* Synchronise `stream` and report whether a persistent kernel of this
 * workspace stopped on its progress watchdog since the last check (the
 * replacement for the reference's expect()-panics on solver failure,
 * sparse/cusparse.rs). */
ddrs_status ddrs_workspace_check(ddrs_workspace_t ws, void* stream);

/* Bind the per-reach static inputs of every following step/window call
 * (device pointers, reference order, length n): Manning n, q_spatial,
 * p_spatial (physical units, i.e. after denormalize, routing/utils.rs:32-41),
 * channel length [m], raw slope (clamped here to params.slope_lb) and
 * x_storage (read only when ddr_match). Copies into schedule order. */
ddrs_status ddrs_mc_bind_statics(ddrs_workspace_t ws, const ddrs_params* params,
                                 const float* d_n, const float* d_q_spatial,
                                 const float* d_p_spatial, const float* d_length,
                                 const float* d_slope, const float* d_x_storage,
                                 void* stream);

/* ---------------------------------------------------------------------------
 * One timestep — the TimestepOp forward (timestep_forward, mmc_op.rs:1704-1857;
 * forward_chain_inner S1..S28, 1061-1336) with the statics bound above.
 * q'_t must already be clamped by the caller (mmc.rs:526). Writes
 * q_next = max(x_sol, discharge_lb) and, if d_x_sol != NULL, the pre-clamp solve.
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_mc_step_fwd(ddrs_workspace_t ws, const ddrs_params* params,
                             const float* d_q_t, const float* d_q_prime_t, float* d_q_next,
                             float* d_x_sol, void* stream);
/* TimestepOp::backward (mmc_op.rs:850-893 / timestep_backward_core 219-848):
 * from the step's inputs, its pre-clamp solve and dL/dQ_{t+1}, writes the five
 * parent gradients (any output may be NULL). Intermediates are recomputed, not
 * saved. */
ddrs_status ddrs_mc_step_bwd(ddrs_workspace_t ws, const ddrs_params* params,
                             const float* d_q_t, const float* d_q_prime_t,
                             const float* d_x_sol, const float* d_grad_q_next,
                             float* d_grad_n, float* d_grad_q_spatial,
                             float* d_grad_p_spatial, float* d_grad_q_t,
                             float* d_grad_q_prime_t, void* stream);

/* ---------------------------------------------------------------------------
 * Window — MuskingumCunge::forward's time loop (mmc.rs:519-632) after
 * setup_inputs: Q[0] = max(Q0, lb); Q[t] = step(Q[t-1], max(q'[t-1], lb)) for
 * t = 1..T-1 (the last forcing row is unused, as in the reference).
 *   d_q_prime  [T][n] raw lateral inflow
 *   d_q0       [n]    initial state (e.g. from ddrs_hotstart)
 *   d_Q_out    [T][n] routed discharge (NULL: not written)
 * Checkpoints Q every `window` steps; keeps the last window's solve history for
 * the adjoint.
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_mc_window_fwd(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                               const float* d_q_prime, const float* d_q0, float* d_Q_out,
                               ddrs_form form, void* stream);
/* Reverse-time adjoint of the last ddrs_mc_window_fwd for the loss whose
 * gradient with respect to Q is d_grad_Q [T][n] (NULL: zero). Recomputes each
 * earlier checkpoint segment's forward before its adjoint. Writes
 *   d_grad_n, d_grad_q_spatial, d_grad_p_spatial [n]  (summed over all steps)
 *   d_grad_q_prime [T][n]  gradient w.r.t. the raw q' through its clamp (NULL: skipped)
 *   d_grad_q0 [n]          dL/dQ0 through the initial clamp (NULL: skipped)
 * Any output may be NULL. */
ddrs_status ddrs_mc_window_adj(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                               const float* d_q_prime, const float* d_q0,
                               const float* d_grad_Q, float* d_grad_n,
                               float* d_grad_q_spatial, float* d_grad_p_spatial,
                               float* d_grad_q_prime, float* d_grad_q0, ddrs_form form,
                               void* stream);

/* Gauge readout (no reference counterpart as an entry point; it fuses the
 * per-gauge select of scatter_add_by_group, src/training/forward.rs:18-34,
 * into the window). After ddrs_workspace_set_gauges(ws, K, h_reaches) (host
 * array of K distinct reach ids, reference order; K = 0 clears), the _gauged
 * window calls read and write only the gauge reaches:
 *   ddrs_mc_window_fwd_gauged writes d_Qg_out [T][K] = Q[t][h_reaches[k]]
 *     (bit-identical to the dense call's Q at those reaches) instead of Q [T][n];
 *   ddrs_mc_window_adj_gauged takes dL/dQ at the gauges, d_grad_Qg [T][K], and
 *     returns what ddrs_mc_window_adj returns for the dense dL/dQ that is zero
 *     elsewhere (bit-identical).
 * The full discharge is then neither written by the forward nor permuted by
 * the adjoint. The reach list keeps its device buffer while later lists fit
 * in it and moves to a larger one (the old one freed) when a list outgrows
 * it, so a HIP graph captured from _gauged calls is valid only until the next
 * ddrs_workspace_set_gauges on that workspace. */
ddrs_status ddrs_workspace_set_gauges(ddrs_workspace_t ws, int32_t K, const int32_t* h_reaches);
ddrs_status ddrs_mc_window_fwd_gauged(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                                      const float* d_q_prime, const float* d_q0, float* d_Qg_out,
                                      ddrs_form form, void* stream);
ddrs_status ddrs_mc_window_adj_gauged(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                                      const float* d_q_prime, const float* d_q0,
                                      const float* d_grad_Qg, float* d_grad_n,
                                      float* d_grad_q_spatial, float* d_grad_p_spatial,
                                      float* d_grad_q_prime, float* d_grad_q0, ddrs_form form,
                                      void* stream);

/* ---------------------------------------------------------------------------
 * Sparse seam: triangular_csr_solve (sparse/mod.rs:483-531) forward and the
 * CsrSolveOp backward (425-472; dispatch.rs:102-138, 214-244) for arbitrary
 * A values in CsrPattern order (nnz).
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_tri_solve_fwd(ddrs_workspace_t ws, const float* d_a_values, const float* d_b,
                               float* d_x, void* stream);
/* grad_b = A^-T grad_x; grad_a[k] = -grad_b[row k] * x[col k] (NULL: skipped). */
ddrs_status ddrs_tri_solve_adj(ddrs_workspace_t ws, const float* d_a_values, const float* d_x,
                               const float* d_grad_x, float* d_grad_b, float* d_grad_a,
                               void* stream);
/* Cold start (mmc.rs:326-361): Q0 = max(solve(I - N, q'_0), discharge_lb). */
ddrs_status ddrs_hotstart(ddrs_workspace_t ws, const ddrs_params* params,
                          const float* d_q_prime0, float* d_q0, void* stream);
/* Its backward (the CsrSolveOp of the cold-start solve through the clamp):
 * grad_q'_0 = (I - N)^-T (grad_Q0 masked where Q0 <= discharge_lb). */
ddrs_status ddrs_hotstart_adj(ddrs_workspace_t ws, const ddrs_params* params,
                              const float* d_q0, const float* d_grad_q0,
                              float* d_grad_q_prime0, void* stream);
/* ddrs_hotstart that allocates nothing (no counterpart in the reference): for
 * callers that cannot call hipMalloc, such as code under HIP graph capture or
 * a serving loop. The same kernel and results as ddrs_hotstart, with its
 * scratch (chunk claims, launch state, one tagged word per schedule position)
 * in a file-scope device array of the library, one per GPU, for plans of at
 * most 4,194,304 schedule positions (32 MiB); a larger plan is rejected with
 * DDRS_ERR_CAPACITY. The plan's schedule must already be on the device (a
 * workspace of the plan created once, at set-up, uploads it) and on the
 * current GPU. *d_err (caller-owned device int) is set to 0, and to 1 when
 * the solve's inputs did not arrive within the watchdog time. Every call on a
 * GPU shares that GPU's scratch: calls that overlap on one GPU (different
 * streams) race; serialise them. */
ddrs_status ddrs_hotstart_noalloc(ddrs_plan_t plan, const ddrs_params* params,
                                  const float* d_q_prime0, float* d_q0, int32_t* d_err,
                                  void* stream);

/* ---------------------------------------------------------------------------
 * Timing of the window kernels of the last window calls on their own stream
 * (HIP events around each launch): number of launches and summed milliseconds
 * of the forward-window kernel and of the adjoint's three kernels
 * (linearisation, reverse-time sweep, gradient finish). ddrs_set_timing(ws, 1)
 * times the last forward and the last adjoint window call; mode 2 sums every
 * call until the next ddrs_set_timing (no host synchronisation until the
 * times are read); 0 turns timing off. No reference counterpart (the
 * reference times with host clocks in benches/).
 * ------------------------------------------------------------------------- */
ddrs_status ddrs_set_timing(ddrs_workspace_t ws, int32_t enable);
ddrs_status ddrs_get_timing(ddrs_workspace_t ws, int32_t* fwd_launches, float* fwd_ms,
                            int32_t* adj_launches, float* adj_ms);
ddrs_status ddrs_get_adj_phase_timing(ddrs_workspace_t ws, float* lin_ms, float* sweep_ms,
                                      float* fin_ms);
/* Launch shape of the persistent kernels: reaches per lane (chunk width / 64)
 * and workgroups (= chunks), forward (default physics) and adjoint sweep. The
 * workspace picks the narrowest width whose chunks fit on the GPU at once;
 * set_chunk_width overrides it (0 keeps the current choice; results do not
 * depend on the width). No reference counterpart. */
ddrs_status ddrs_workspace_layout(ddrs_workspace_t ws, int32_t* rpl_fwd, int32_t* grid_fwd,
                                  int32_t* rpl_sweep, int32_t* grid_sweep);
ddrs_status ddrs_workspace_set_chunk_width(ddrs_workspace_t ws, int32_t rpl_fwd,
                                           int32_t rpl_sweep);
/* Linearise the adjoint while the forward window runs (enable = 1; off by
 * default). A persistent ddrs_mc_window_fwd then also computes, on a second
 * stream of the workspace (default priority) ordered against `stream` by
 * events, the per-step coefficients and local backward that the following
 * ddrs_mc_window_adj would compute first, reading the forward's history as it
 * is published; ddrs_mc_window_adj with the same parameters and q' skips that
 * phase. Worth enabling only when an adjoint follows every forward. grid: the
 * one-wave workgroups of the concurrent launch (0: 16 per compute unit). Results
 * are bit-identical either way. No reference counterpart (the reference's
 * backward in src/routing/mmc_op.rs:1704-1857 runs after the forward).
 * Currently disabled: enable = 1 returns DDRS_ERR_INVALID (a 350000-reach,
 * 720-step window with the concurrent launch did not finish; the cause is not
 * known); enable = 0 is accepted. */
ddrs_status ddrs_workspace_set_linearize_in_forward(ddrs_workspace_t ws, int32_t enable,
                                                   int32_t grid);
/* Per-chunk trace of the persistent forward (which = 0) and adjoint-sweep
 * (which = 1) launches: for each chunk, 8 values: GPU wall clock (100 MHz
 * ticks) when it was claimed, when its first and last steps were solved; the
 * number of polls that found an input missing; and ticks spent in prepare, in
 * the first poll of each step, waiting for late inputs, and in solve. h_rec
 * holds cap chunk records (8 each); n_chunks gets the chunk count of the last
 * launch. which = 2: the items (64 positions x 32 steps) of the last
 * linearisation run beside a forward window, 8 values each: wall clock when
 * taken and when done, and 1 if the concurrent launch did it (0: the launch
 * after the forward); then one record whose first value counts the concurrent
 * launch's workgroups that left without an item. No reference counterpart. */
ddrs_status ddrs_set_trace(ddrs_workspace_t ws, int32_t enable);
ddrs_status ddrs_get_trace(ddrs_workspace_t ws, int32_t which, uint64_t* h_rec, int32_t cap,
                           int32_t* n_chunks);

/* ---------------------------------------------------------------------------
 * Index construction (host, bit-exact): the reference's topological order
 * (adjacency/build.rs:170-272; rustworkx/petgraph DFS finish-time rule) and
 * per-batch compression (data/collate.rs:197-349).
 * ------------------------------------------------------------------------- */
/* records: n_records rows of (comid, up1, up2, up3, up4), zero = no upstream.
 * h_order gets the COMIDs in topological order (capacity order_cap), h_rows /
 * h_cols the lower-triangular COO (capacity edge_cap). Counts are returned. */
ddrs_status ddrs_build_topology(int32_t n_records, const int64_t* h_records,
                                int64_t* h_order, int32_t order_cap, int32_t* n_order,
                                int32_t* h_rows, int32_t* h_cols, int32_t edge_cap,
                                int32_t* n_edges, int64_t* h_dropped, int32_t dropped_cap,
                                int32_t* n_dropped);
/* Unioned COO (CONUS positions) + gauge positions -> compressed positions.
 * h_active (capacity active_cap) gets the sorted active set; h_rows/h_cols the
 * compressed edges (n_edges entries, in sorted (row, col) order after dedup);
 * h_gauge the compressed gauge positions. */
ddrs_status ddrs_compress(int32_t n_edges_in, const int32_t* h_rows_in,
                          const int32_t* h_cols_in, int32_t n_gauges,
                          const int32_t* h_gauges, int32_t* h_active, int32_t active_cap,
                          int32_t* n_active, int32_t* h_rows, int32_t* h_cols,
                          int32_t* n_edges, int32_t* h_gauge);
/* ddrs_compress with the rest of collate::compress (data/collate.rs:197-349):
 * - h_conus_parent_offset (n_conus_parent + 1 entries, n_conus_parent = 0: no
 *   map) is the reach-subdivision map in CONUS sub-reach rows; h_parent_offset
 *   (capacity active_cap + 1) gets it in compressed rows (*n_parent parents,
 *   *n_parent + 1 entries). A parent present only in part is an error
 *   (compressed_parent_offset, 141-180).
 * - outflow_idx (297-336), as CSR: gauge g's prediction sums the discharge of
 *   compressed reaches h_outflow[h_outflow_ptr[g] .. h_outflow_ptr[g+1]]
 *   (h_outflow_ptr has n_gauges + 1 entries; h_outflow capacity outflow_cap).
 *   ddr_match != 0: the gauge's upstream columns in edge order, or the gauge
 *   itself when it has none; ddr_match == 0: the gauge's own reach, at the
 *   last piece of its parent when a map is given. */
ddrs_status ddrs_compress_ex(int32_t n_edges_in, const int32_t* h_rows_in,
                             const int32_t* h_cols_in, int32_t n_gauges,
                             const int32_t* h_gauges, int32_t ddr_match,
                             int32_t n_conus_parent, const int32_t* h_conus_parent_offset,
                             int32_t* h_active, int32_t active_cap, int32_t* n_active,
                             int32_t* h_rows, int32_t* h_cols, int32_t* n_edges,
                             int32_t* h_gauge, int32_t* h_parent_offset, int32_t* n_parent,
                             int32_t* h_outflow_ptr, int32_t* h_outflow, int32_t outflow_cap,
                             int32_t* n_outflow);

/* ---------------------------------------------------------------------------
 * KAN parameter head (device, f32): per-reach attributes -> routing parameters.
 * Replaces KanHead::forward (src/nn/kan_head.rs:215-246: input Linear, the
 * KanLayer blocks of rskan v0.1.3 / pykan, output Linear, sigmoid, [P, N]) and
 * its autodiff backward. All pointers are device pointers; weight matrices are
 * [out][in] row-major (the fixture's layout, src/nn/kan_head.rs:299-360).
 * ------------------------------------------------------------------------- */
typedef struct {
    int32_t F; /* input attributes (<= 63) */
    int32_t H; /* hidden width (<= 32) */
    int32_t P; /* output parameters (<= 16) */
    int32_t L; /* KanLayer(H, H) blocks (<= 8) */
    int32_t G; /* knots per input: grid + 2k + 1 */
    int32_t k; /* spline order (<= 5); coefficients per edge C = G - 1 - k <= 63 */
} ddrs_kan_dims;

typedef struct {
    const float* input_weight;  /* [H][F] */
    const float* input_bias;    /* [H] */
    const float* grid;          /* [L][H][G] knots, ascending per input */
    const float* coef;          /* [L][H in][H out][C] */
    const float* scale_base;    /* [L][H in][H out] */
    const float* scale_sp;      /* [L][H in][H out] */
    const float* mask;          /* [L][H in][H out] */
    const float* output_weight; /* [P][H] */
    const float* output_bias;   /* [P] */
} ddrs_kan_params;

/* Gradients of the trained parameters (grid and mask are not trained), same
 * layouts; overwritten by ddrs_kan_head_bwd. */
typedef struct {
    float* input_weight;
    float* input_bias;
    float* coef;
    float* scale_base;
    float* scale_sp;
    float* output_weight;
    float* output_bias;
} ddrs_kan_grads;

/* Workspace the calls below need for n reaches (device memory the caller
 * allocates; its contents need not persist between calls). No reference
 * counterpart (burn allocates inside its ops). */
ddrs_status ddrs_kan_workspace_bytes(const ddrs_kan_dims* dims, int32_t n, size_t* fwd_bytes,
                                     size_t* bwd_bytes);
/* d_out[p][r] = sigmoid(head(d_inputs[r][:]))[p] for r < n (reference
 * KanHead::forward, src/nn/kan_head.rs:215). */
ddrs_status ddrs_kan_head_fwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, float* d_out, void* d_workspace,
                              size_t workspace_bytes, void* stream);
/* Reverse mode of ddrs_kan_head_fwd for d_grad_out [P][n]: every parameter
 * gradient into `grads`, and dL/d inputs into d_grad_inputs [n][F] unless it is
 * NULL (burn's autodiff of the same graph in the reference). Deterministic:
 * the sums over reaches run in a fixed order. */
ddrs_status ddrs_kan_head_bwd(const ddrs_kan_dims* dims, const ddrs_kan_params* params,
                              int32_t n, const float* d_inputs, const float* d_grad_out,
                              const ddrs_kan_grads* grads, float* d_grad_inputs,
                              void* d_workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Gauge readout and training loss (device, f32, deterministic).
 * Replaces scatter_add_by_group (src/training/forward.rs:18-34),
 * tau_trim_and_downsample (src/training/loss.rs:33-95), nse_batch_loss
 * (src/training/loss.rs:226-237) and the micro-batch objective of
 * src/training/driver.rs:110-201 (warmup slice, NaN-gauge filter). A discharge
 * series is addressed by strides: element (reach n, hour t) at
 * d_runoff[n * stride_n + t * stride_t]; the reference's (N, T) is
 * (stride_n, stride_t) = (T, 1), this library's time-major [T][N] is (1, N).
 * ------------------------------------------------------------------------- */
typedef struct ddrs_gauge_plan* ddrs_gauge_plan_t;
/* The outflow map of a batch: entry k adds reach h_flat_indices[k] into gauge
 * h_group_ids[k] (RoutingTensors::flat_indices / group_ids, data/dataset.rs:
 * 139-141, built from outflow_idx at :160-230). Host arrays, any order; sums
 * run in increasing k as the reference's scatter does. Scratch for series of
 * up to max_t_hours hours is allocated here, none in the calls below. At most
 * 65535 gauges and 65535 distinct outflow reaches. */
ddrs_status ddrs_gauge_plan_create(int32_t n_reaches, int32_t num_gauges, int32_t K,
                                   const int32_t* h_flat_indices, const int32_t* h_group_ids,
                                   int32_t max_t_hours, ddrs_gauge_plan_t* out);
void ddrs_gauge_plan_destroy(ddrs_gauge_plan_t plan);
/* d_out[g][t] = sum over entries k of gauge g of runoff(flat_k, t), t < T. */
ddrs_status ddrs_scatter_add_by_group(ddrs_gauge_plan_t plan, const float* d_runoff, int32_t T,
                                      int64_t stride_n, int64_t stride_t, float* d_out,
                                      void* stream);
/* Backward: grunoff(n, t) = sum over entries k with flat_k = n of d_gout[group_k][t], written
 * for every reach n of the map (the scatter's select backward); no other entry is written. */
ddrs_status ddrs_scatter_add_by_group_bwd(ddrs_gauge_plan_t plan, const float* d_gout,
                                          int32_t T, int64_t stride_n, int64_t stride_t,
                                          float* d_grunoff, void* stream);
/* d_out [G][T_days] = hours [tau, T_hours - (24 - tau)) of d_in [G][T_hours] pooled with the
 * area weights of area_pool_weights; T_days = (T_hours - 24) / 24; 0 <= tau < 24. */
ddrs_status ddrs_tau_trim_downsample(const float* d_in, int32_t G, int32_t T_hours, int32_t tau,
                                     float* d_out, void* stream);
/* Backward into every entry of d_gin [G][T_hours] (zero outside the window). */
ddrs_status ddrs_tau_trim_downsample_bwd(const float* d_gout, int32_t G, int32_t T_hours,
                                         int32_t tau, float* d_gin, void* stream);
/* *d_loss = mean over [G][T] of (p - o)^2 / (sigma_g + eps)^2 (G <= the plan's gauges; the
 * plan lends its scratch). Backward: d_gp = gloss * dloss/dp, gloss = *d_gloss or 1. */
ddrs_status ddrs_nse_batch_loss(ddrs_gauge_plan_t plan, const float* d_p, const float* d_o,
                                const float* d_sigma, int32_t G, int32_t T, float eps,
                                float* d_loss, void* stream);
ddrs_status ddrs_nse_batch_loss_bwd(const float* d_p, const float* d_o, const float* d_sigma,
                                    int32_t G, int32_t T, float eps, const float* d_gloss,
                                    float* d_gp, void* stream);
/* The objective of one micro-batch with loss.kind nse-batch (driver.rs:110-201):
 * gauge readout -> tau trim + daily pooling -> days [warmup, T_days) -> gauges with any NaN
 * in d_obs [G][T_days] there dropped -> mean of (p - o)^2 / (sigma + eps)^2. Writes *d_loss
 * (0 when no gauge is kept), *d_n_valid (kept gauges x days, loss_denominator, if not NULL),
 * d_hourly [G][T_hours] (the readout, if not NULL) and, if d_grunoff is not NULL, dL/d runoff
 * at the map's reaches (other entries not written). */
ddrs_status ddrs_gauge_nse_loss(ddrs_gauge_plan_t plan, const float* d_runoff, int32_t T_hours,
                                int64_t stride_n, int64_t stride_t, int32_t tau, int32_t warmup,
                                const float* d_obs, const float* d_sigma, float eps,
                                float* d_loss, int32_t* d_n_valid, float* d_grunoff,
                                float* d_hourly, void* stream);
/* The same objective for every loss kind of batch_loss (src/training/loss.rs:179-208;
 * LossKind and LossConfig, src/config.rs:220-289):
 *   DDRS_LOSS_L1         the mean of |p - o| (loss.rs:186; gradient sign(p - o) / n_valid,
 *                        0 where p == o; d_sigma unused and may be NULL);
 *   DDRS_LOSS_NSE_BATCH  as ddrs_gauge_nse_loss;
 *   DDRS_LOSS_NNSE_KGE   per kept gauge nnse_weight (1 - NNSE) + kge_weight (1 - KGE),
 *                        averaged over gauges (loss.rs:257-318);
 *   DDRS_LOSS_KGE        per kept gauge min(r_w (r-1)^2 + alpha_w (alpha-1)^2 +
 *                        beta_w (beta-1)^2, kge_clamp) + nnse_weight (1 - NNSE), averaged
 *                        over gauges (loss.rs:320-373; the clamp passes gradient where the
 *                        sum is at most kge_clamp).
 * *d_n_valid is loss_denominator (loss.rs:164-169): kept gauges x days for the first two,
 * kept gauges for the per-gauge kinds. Where the reference's gradient is NaN (1 - KGE = 0
 * in NNSE_KGE), so is this one. */
enum { DDRS_LOSS_L1 = 0, DDRS_LOSS_NSE_BATCH = 1, DDRS_LOSS_NNSE_KGE = 2, DDRS_LOSS_KGE = 3 };
typedef struct ddrs_loss_config {
    int32_t kind;       /* DDRS_LOSS_* */
    float eps;          /* LossConfig.eps (0.1) */
    float nnse_weight;  /* 1 */
    float kge_weight;   /* 1 */
    float r_weight;     /* 1 */
    float alpha_weight; /* 1 */
    float beta_weight;  /* 1 */
    float kge_clamp;    /* 10 */
} ddrs_loss_config;
ddrs_status ddrs_gauge_batch_loss(ddrs_gauge_plan_t plan, const ddrs_loss_config* cfg,
                                  const float* d_runoff, int32_t T_hours, int64_t stride_n,
                                  int64_t stride_t, int32_t tau, int32_t warmup,
                                  const float* d_obs, const float* d_sigma, float* d_loss,
                                  int32_t* d_n_valid, float* d_grunoff, float* d_hourly,
                                  void* stream);

/* ---------------------------------------------------------------------------
 * Network stores on disk (host): zarr v3 arrays and group attributes, as the
 * reference reads its binsparse-COO stores with the zarrs crate
 * (src/data/store/zarr.rs:47-300: order, indices_0/1, length_m, slope,
 * parent_order / parent_offset; gage_idx / gage_catchment attributes).
 * Supported: regular chunk grids (C order), "bytes" (little or big endian) +
 * optional "zstd" (decoded by the system's libzstd.so.1), default or v2 chunk
 * keys; a missing chunk reads as the fill value. Anything else is an error.
 * ------------------------------------------------------------------------- */
typedef enum {
    DDRS_ZARR_I8 = 1, DDRS_ZARR_U8 = 2, DDRS_ZARR_I16 = 3, DDRS_ZARR_U16 = 4,
    DDRS_ZARR_I32 = 5, DDRS_ZARR_U32 = 6, DDRS_ZARR_I64 = 7, DDRS_ZARR_U64 = 8,
    DDRS_ZARR_F32 = 9, DDRS_ZARR_F64 = 10
} ddrs_zarr_dtype;
/* Data type, rank, shape (first shape_cap dims) and element count of the
 * array stored in directory array_dir (its zarr.json). Any output may be NULL. */
ddrs_status ddrs_zarr_array_info(const char* array_dir, int32_t* dtype, int32_t* ndim,
                                 int64_t* shape, int32_t shape_cap, int64_t* count);
/* The whole array, C order, into h_out (capacity elements); dtype must be the
 * stored type (no conversion, as the reference's typed reads). */
ddrs_status ddrs_zarr_read_array(const char* array_dir, int32_t dtype, void* h_out,
                                 int64_t capacity);
/* Attribute `key` of the group in group_dir: its number (NaN if not numeric)
 * and its text (strings, or a number's JSON text), either output optional. */
ddrs_status ddrs_zarr_group_attr(const char* group_dir, const char* key, double* h_num,
                                 char* h_text, int32_t text_cap);

/* ---------------------------------------------------------------------------
 * Synthetic workloads (host), seeded and portable.
 * ------------------------------------------------------------------------- */
/* Reach i < n-1 drains to a reach drawn uniformly from (i, min(i+window, n-1)];
 * writes n-1 edges (rows = downstream, cols = upstream). */
ddrs_status ddrs_gen_tree(int32_t n, int32_t window, uint64_t seed, int32_t* h_rows,
                          int32_t* h_cols);
/* length U[1,10] km (m), slope log-U[1e-4, 5e-2], n U[0.015, 0.25], q_spatial U[0,1]. */
ddrs_status ddrs_gen_attributes(int32_t n, uint64_t seed, float* h_length, float* h_slope,
                                float* h_n, float* h_q_spatial);
/* q'[T][n] ~ LogNormal(ln 0.5, 1.0); if storms != 0, times (1 + m_t) where m_t
 * is a sum of Poisson(1/72 h) Gamma(2, 2) pulses decaying with e^{-1/24} per h.
 * Device variant writes a device buffer on `stream`; host variant a host buffer. */
ddrs_status ddrs_gen_forcing(int32_t n, int32_t T, uint64_t seed, int32_t storms,
                             float* h_q_prime);
ddrs_status ddrs_gen_forcing_device(int32_t n, int32_t T, uint64_t seed, int32_t storms,
                                    float* d_q_prime, void* stream);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference ddrs CPU path for the differentiable
 * Muskingum-Cunge timestep, used ONLY by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker. The product (ddrs_amd/, the HIP
 * library) never links, loads or calls anything here.
 *
 * Every routine exists three times: an `_f` variant computing in float (the
 * reference's arithmetic type, op-for-op in the reference's order, with
 * glibc's powf / logf as the reference's libm calls), a `_d` variant computing
 * the same algorithm in double (the FP64 oracle used for finite-difference
 * gradient checks and for stating FP64 error), and a `_c` variant: `_f` with
 * the device's correctly rounded fp32 power and logarithm (devmath.cpp), which
 * separates the device kernels' arithmetic from libm's rounding.
 *
 * Reference files (taddyb/ddrs):
 *   src/routing/mmc_op.rs   forward_chain_inner 1061-1336, timestep_backward_core 219-848
 *   src/routing/mmc.rs      setup_inputs 239-376, forward 519-632, DT_SECONDS 33
 *   src/sparse/mod.rs       from_sparse 173-227, build_transposed_pattern 276-304,
 *                           forward_sub_lower 314-335, back_sub_upper_transposed 341-367,
 *                           assemble_primitive 652-723, cpu_spmv_forward 731-749,
 *                           cpu_assemble_backward 752-770, cpu_spmv_backward 773-792
 *   src/routing/utils.rs    denormalize 32-41
 *
 * Parity pins (see tests/): GOLDEN_X_SOL_BITS of tests/positivity_clamp.rs
 * (bit-exact S1..S27 in f32), the hotstart_77006074 npz fixture, closed forms
 * (chain cumsum, c1+c2+c3=1, beta limits, Cunge-X identity), mass
 * conservation and FP64 central finite differences for every gradient.
 */
#ifndef DDRS_MC_ORACLE_H
#define DDRS_MC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar routing parameters (reference: config.rs AttributeMinimums 426-446,
 * Params.ddr_match / enforce_positivity 580-672, mmc.rs DT_SECONDS 33). */
typedef struct {
    double dt;
    double discharge_lb;
    double depth_lb;
    double velocity_lb;
    double bottom_width_lb;
    double slope_lb;
    int ddr_match;
    int enforce_positivity;
} orc_params;

/* Indices of the per-reach saved intermediates of one forward step. The
 * layout is [ORC_NSAVE][n]; it carries everything timestep_backward_core reads
 * (TimestepState, mmc_op.rs:103-163). */
enum {
    ORC_Q_EPS = 0, ORC_RATIO, ORC_DENOMINATOR, ORC_DEPTH, ORC_TOP_WIDTH,
    ORC_SIDE_SLOPE_RAW, ORC_SIDE_SLOPE, ORC_BW_RAW, ORC_BOTTOM_WIDTH,
    ORC_HYD_RADIUS, ORC_V_UN, ORC_V_CL, ORC_CELERITY, ORC_K_MUSK, ORC_X_EFF,
    ORC_DENOM, ORC_C1, ORC_C2, ORC_C3, ORC_C4, ORC_I_T, ORC_B_RHS, ORC_X_SOL,
    ORC_NSAVE
};

/* Threads of the per-reach loops of the step functions (1, the default:
 * serial; results do not depend on it). For the CPU baseline's parallel leg. */
extern int orc_threads;
void orc_set_threads(int n);

/* ---- CsrPattern construction (sparse/mod.rs:173-304). Integer work, one
 * variant. Outputs sized n+1 / nnz_off+n as in the reference. Returns 0, or
 * -1 if an entry is not strictly lower triangular. */
int orc_pattern_from_coo(int n, int nnz_off, const int* rows, const int* cols,
                         const float* vals, int* crow, int* col, int* row_for_nnz,
                         float* adj_values, float* diag_mask, int* trans_crow,
                         int* trans_col, int* trans_to_orig);

#define ORC_DECLARE(SUF, REAL)                                                        \
    void orc_forward_sub_lower_##SUF(int n, const int* crow, const int* col,         \
                                     const REAL* a, const REAL* b, REAL* x);          \
    void orc_back_sub_upper_transposed_##SUF(int n, const int* trans_crow,           \
                                             const int* trans_col,                    \
                                             const int* trans_to_orig, const REAL* a, \
                                             const REAL* b, REAL* y);                 \
    void orc_spmv_##SUF(int n, int nnz, const int* row_for_nnz, const int* col,      \
                        const REAL* adj, const REAL* q, REAL* out);                  \
    void orc_assemble_##SUF(int nnz, const int* row_for_nnz, const REAL* adj,        \
                            const REAL* diag_mask, const REAL* c, REAL* a);           \
    void orc_denormalize_##SUF(int n, const REAL* v, double lo, double hi,           \
                               int log_space, REAL* out);                             \
    void orc_step_fwd_##SUF(const orc_params* P, int n, int nnz, const int* crow,    \
                            const int* col, const int* row_for_nnz, const REAL* adj, \
                            const REAL* diag_mask, const REAL* nman, const REAL* qsp, \
                            const REAL* psp, const REAL* q_t, const REAL* qp_t,       \
                            const REAL* length, const REAL* slope, const REAL* xst,   \
                            REAL* saved, REAL* a_values, REAL* q_next);               \
    void orc_step_bwd_##SUF(const orc_params* P, int n, int nnz, const int* col,     \
                            const int* row_for_nnz, const REAL* adj,                  \
                            const int* trans_crow, const int* trans_col,              \
                            const int* trans_to_orig, const REAL* nman,               \
                            const REAL* psp, const REAL* q_t, const REAL* qp_t,       \
                            const REAL* length, const REAL* slope,                    \
                            const REAL* saved, const REAL* a_values, const REAL* gy,  \
                            REAL* gn, REAL* gqsp, REAL* gp, REAL* gq, REAL* gqp);     \
    void orc_hotstart_##SUF(const orc_params* P, int n, int nnz, const int* crow,    \
                            const int* col, const int* row_for_nnz, const REAL* adj, \
                            const REAL* diag_mask, const REAL* qp0, REAL* q0);        \
    void orc_route_fwd_##SUF(const orc_params* P, int n, int nnz, int T,             \
                             const int* crow, const int* col, const int* row_for_nnz, \
                             const REAL* adj, const REAL* diag_mask, const REAL* nman,\
                             const REAL* qsp, const REAL* psp, const REAL* length,    \
                             const REAL* slope_raw, const REAL* xst,                  \
                             const REAL* qprime_raw, const REAL* q0_in, REAL* Q,      \
                             REAL* X);                                                \
    int orc_route_adj_##SUF(const orc_params* P, int n, int nnz, int T,              \
                            const int* crow, const int* col, const int* row_for_nnz,  \
                            const REAL* adj, const REAL* diag_mask,                   \
                            const int* trans_crow, const int* trans_col,              \
                            const int* trans_to_orig, const REAL* nman,               \
                            const REAL* qsp, const REAL* psp, const REAL* length,     \
                            const REAL* slope_raw, const REAL* xst,                   \
                            const REAL* qprime_raw, const REAL* q0_in,                \
                            const REAL* gQ, REAL* gn, REAL* gqsp, REAL* gp,           \
                            REAL* gqprime, REAL* gq0, REAL* Q_out, REAL* X_out);      \
    int orc_route_path_##SUF(const orc_params* P, int n, int T, const int* crow,     \
                             const int* col, const REAL* adj, const REAL* diag_mask,  \
                             const REAL* nman, const REAL* qsp, const REAL* psp,      \
                             const REAL* length, const REAL* slope_raw,               \
                             const REAL* xst, const REAL* qprime_raw, const REAL* Qb, \
                             const REAL* Xb, int m, const int* path, REAL* Qp);

ORC_DECLARE(f, float)
ORC_DECLARE(d, double)
/* float with the device's powers and logarithms (devmath.cpp) */
ORC_DECLARE(c, float)

#ifdef __cplusplus
}
#endif
#endif

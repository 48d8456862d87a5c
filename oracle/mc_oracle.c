/*
 * CPU ORACLE — test infrastructure only (see mc_oracle.h). Instantiates the
 * REAL-generic body for float and double and holds the integer CsrPattern
 * construction.
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the per-reach loops of orc_step_fwd / orc_step_bwd (1: serial,
 * the default). Each reach's values there are computed independently, so the
 * results do not depend on it; only the CPU baseline's parallel leg sets it. */
int orc_threads = 1;
void orc_set_threads(int n) { orc_threads = (n > 1) ? n : 1; }

/* powf_scalar(2.0) of the reference (S13, S17) is the exact square rounded
 * once, written x * x. (gcc emits x * x for powf(x, 2.0f) as well; a run-time
 * glibc powf(x, 2) is not correctly rounded on ~0.07% of x in [0.5, 50].) */

/* ---------------- float instantiation ---------------- */
#define REAL float
#define SUF f
#define RPow powf
#define RSqrt sqrtf
#define RLog logf
#define RExp expf
#define RSq(x) ((x) * (x))
#include "mc_oracle_body.inc"
#undef REAL
#undef SUF
#undef RPow
#undef RSqrt
#undef RLog
#undef RExp
#undef RSq

/* ---------------- double instantiation ---------------- */
#define REAL double
#define SUF d
#define RPow pow
#define RSqrt sqrt
#define RLog log
#define RExp exp
#define RSq(x) ((x) * (x))
#include "mc_oracle_body.inc"
#undef REAL
#undef SUF
#undef RPow
#undef RSqrt
#undef RLog
#undef RExp
#undef RSq

/* ---------------- float with the device's powers (devmath.cpp) ----------------
 * The same float arithmetic with the device's correctly rounded fp32 power
 * and logarithm in place of glibc's powf / logf (everything else, sqrtf and
 * expf included, unchanged): the _c variant. */
float orc_devpowf(float x, float y);
float orc_devlogf(float x);
#define REAL float
#define SUF c
#define RPow orc_devpowf
#define RSqrt sqrtf
#define RLog orc_devlogf
#define RExp expf
#define RSq(x) ((x) * (x))
#include "mc_oracle_body.inc"
#undef REAL
#undef SUF
#undef RPow
#undef RSqrt
#undef RLog
#undef RExp
#undef RSq

/* ---------------- CsrPattern::from_sparse ---------------- */

/* Stable counting sort of idx[0..m) by key[idx[.]] in [0, n) into out. */
static void counting_sort(int m, int n, const int* key, const int* idx, int* out) {
    int* start = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int j = 0; j < m; ++j) start[key[idx[j]] + 1] += 1;
    for (int i = 0; i < n; ++i) start[i + 1] += start[i];
    for (int j = 0; j < m; ++j) out[start[key[idx[j]]]++] = idx[j];
    free(start);
}

/* CsrPattern::from_sparse (sparse/mod.rs:173-227): sort COO by (row, col),
 * emit each row's off-diagonals in ascending column order followed by the
 * diagonal; then build_transposed_pattern (sparse/mod.rs:276-304). */
int orc_pattern_from_coo(int n, int nnz_off, const int* rows, const int* cols,
                         const float* vals, int* crow, int* col, int* row_for_nnz,
                         float* adj_values, float* diag_mask, int* trans_crow,
                         int* trans_col, int* trans_to_orig) {
    for (int k = 0; k < nnz_off; ++k)
        if (!(rows[k] > cols[k]) || rows[k] >= n || cols[k] < 0) return -1;
    /* entries ordered by (row, col, input index): a stable sort by column,
     * then a stable sort by row (no shared state, so threads may build
     * patterns concurrently) */
    int* order = (int*)malloc(sizeof(int) * (size_t)(nnz_off > 0 ? nnz_off : 1));
    int* tmp = (int*)malloc(sizeof(int) * (size_t)(nnz_off > 0 ? nnz_off : 1));
    for (int k = 0; k < nnz_off; ++k) tmp[k] = k;
    counting_sort(nnz_off, n, cols, tmp, order);
    memcpy(tmp, order, sizeof(int) * (size_t)nnz_off);
    counting_sort(nnz_off, n, rows, tmp, order);
    free(tmp);
    int nnz = nnz_off + n;
    int cursor = 0, w = 0;
    crow[0] = 0;
    for (int i = 0; i < n; ++i) {
        while (cursor < nnz_off && rows[order[cursor]] == i) {
            int k = order[cursor];
            col[w] = cols[k];
            row_for_nnz[w] = i;
            adj_values[w] = vals[k];
            diag_mask[w] = 0.0f;
            ++w;
            ++cursor;
        }
        col[w] = i;
        row_for_nnz[w] = i;
        adj_values[w] = 0.0f;
        diag_mask[w] = 1.0f;
        ++w;
        crow[i + 1] = w;
    }
    free(order);
    /* transposed view */
    int* counts = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int k = 0; k < nnz; ++k) counts[col[k] + 1] += 1;
    trans_crow[0] = 0;
    for (int i = 1; i <= n; ++i) trans_crow[i] = trans_crow[i - 1] + counts[i];
    int* cur = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(cur, trans_crow, sizeof(int) * ((size_t)n + 1));
    for (int k = 0; k < nnz; ++k) {
        int r = col[k];
        int pos = cur[r]++;
        trans_col[pos] = row_for_nnz[k];
        trans_to_orig[pos] = k;
    }
    free(counts);
    free(cur);
    return 0;
}

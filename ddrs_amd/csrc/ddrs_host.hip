// Host side of the C ABI (include/ddrs_hip.h): network plan, workspace and the
// step / window / sparse-seam entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ddrs_internal.h"

using namespace ddrs;

namespace {

thread_local std::string g_err;

ddrs_status fail(ddrs_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(DDRS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t upload(T** dptr, const std::vector<T>& h) {
    *dptr = nullptr;
    size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dptr), bytes);
    if (e != hipSuccess) return e;
    if (!h.empty()) e = hipMemcpy(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

struct ddrs_plan {
    int R = 0, E = 0, nnz = 0, C = 0, nlevels = 0;
    // CsrPattern (sparse/mod.rs:115-146)
    std::vector<int> crow, col, row_for_nnz, trans_crow, trans_col, trans_to_orig;
    std::vector<float> adj_values, diag_mask;
    // schedule
    std::vector<int> level, perm;
    std::vector<void*> dev;  // every device allocation, freed on destroy
    // schedule in host memory (uploaded when a workspace is created)
    std::vector<int> h_up_ptr, h_up_pos, h_up_k, h_dn_ptr, h_dn_pos, h_dn_k, h_diag_k, h_nsub,
        h_nsubr;
    std::vector<float> h_up_adj, h_dn_adj;
    std::vector<unsigned char> h_sub, h_subr;
    bool on_device = false;
    Sched sc{};
    int* d_row_for_nnz = nullptr;
    int* d_col = nullptr;
};

struct ddrs_workspace {
    ddrs_plan* plan = nullptr;
    int window = 0, max_steps = 0, max_seg = 0;
    size_t bytes = 0;
    std::vector<void*> dev;
    Statics st{};
    float *st_q = nullptr, *st_i = nullptr, *st_lam = nullptr;
    float *acc_n = nullptr, *acc_qsp = nullptr, *acc_p = nullptr;
    float *sc_q = nullptr, *sc_i = nullptr, *sc_qp = nullptr;
    float* ck = nullptr;  // [max_seg][2][R]
    uint64_t *hist = nullptr, *amsg = nullptr, *slot = nullptr;
    int* err = nullptr;
    uint32_t epoch = 0;
    int grid_fwd[3] = {0, 0, 0}, grid_adj[3] = {0, 0, 0}, grid_tri_fwd = 0, grid_tri_adj = 0;
    bool statics_bound = false;
    // last forward window
    int last_T = -1, hist_seg = -1;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_fwd, ev_adj;
    int n_fwd = 0, n_adj = 0;
};

namespace {

ddrs_status dalloc(ddrs_workspace* ws, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess)
        return fail(DDRS_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                        "): " + hipGetErrorString(e));
    ws->dev.push_back(*p);
    ws->bytes += bytes;
    return DDRS_OK;
}

int variant(const ddrs_params* p) { return p->ddr_match ? 1 : (p->enforce_positivity ? 2 : 0); }

Phys to_phys(const ddrs_params* p) {
    Phys P;
    P.dt = p->dt;
    P.q_lb = p->discharge_lb;
    P.depth_lb = p->depth_lb;
    P.v_lb = p->velocity_lb;
    P.bw_lb = p->bottom_width_lb;
    return P;
}

ddrs_status check_params(const ddrs_params* p) {
    if (!p) return fail(DDRS_ERR_INVALID, "params is NULL");
    if (!(p->dt > 0)) return fail(DDRS_ERR_INVALID, "params.dt must be > 0");
    return DDRS_OK;
}

void record(ddrs_workspace* ws, std::vector<hipEvent_t>& pool, int& n, hipStream_t s) {
    if (!ws->timing) return;
    if (static_cast<int>(pool.size()) <= n) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        pool.push_back(e);
    }
    hipEventRecord(pool[n], s);
    ++n;
}

}  // namespace

extern "C" {

const char* ddrs_last_error(void) { return g_err.c_str(); }
void ddrs_internal_set_error(const char* msg) { g_err = msg ? msg : ""; }
const char* ddrs_version(void) { return "ddrs_hip 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// plan

ddrs_status ddrs_plan_create(int32_t n, int32_t nnz_off, const int32_t* h_rows,
                             const int32_t* h_cols, const float* h_vals, ddrs_plan_t* out) {
    if (!out) return fail(DDRS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n <= 0) return fail(DDRS_ERR_INVALID, "network must have at least one reach");
    if (nnz_off < 0 || (nnz_off > 0 && (!h_rows || !h_cols)))
        return fail(DDRS_ERR_INVALID, "bad COO arrays");
    for (int k = 0; k < nnz_off; ++k) {
        if (h_rows[k] < 0 || h_rows[k] >= n || h_cols[k] < 0 || h_cols[k] >= n)
            return fail(DDRS_ERR_INVALID, "COO index out of range at entry " + std::to_string(k));
        if (!(h_rows[k] > h_cols[k]))
            return fail(DDRS_ERR_INVALID,
                        "adjacency not strictly lower triangular at entry " + std::to_string(k));
    }
    auto* pl = new ddrs_plan();
    pl->R = n;
    pl->E = nnz_off;
    pl->nnz = nnz_off + n;
    const int R = n, nnz = pl->nnz;

    // CsrPattern::from_sparse: entries ordered by (row, col), each row's
    // off-diagonals then its diagonal. Bucket by row, then sort columns within
    // each row (stable, so duplicate edges keep their input order).
    std::vector<int> cnt(R + 1, 0);
    for (int k = 0; k < nnz_off; ++k) cnt[h_rows[k] + 1]++;
    for (int i = 0; i < R; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> byrow(nnz_off);
    {
        std::vector<int> cur(cnt.begin(), cnt.end() - 1);
        for (int k = 0; k < nnz_off; ++k) byrow[cur[h_rows[k]]++] = k;
    }
    pl->crow.assign(R + 1, 0);
    pl->col.resize(nnz);
    pl->row_for_nnz.resize(nnz);
    pl->adj_values.resize(nnz);
    pl->diag_mask.resize(nnz);
    int w = 0;
    for (int i = 0; i < R; ++i) {
        auto b = byrow.begin() + cnt[i], e = byrow.begin() + cnt[i + 1];
        std::stable_sort(b, e, [&](int a, int c) { return h_cols[a] < h_cols[c]; });
        for (auto it = b; it != e; ++it) {
            pl->col[w] = h_cols[*it];
            pl->row_for_nnz[w] = i;
            pl->adj_values[w] = h_vals ? h_vals[*it] : 1.0f;
            pl->diag_mask[w] = 0.0f;
            ++w;
        }
        pl->col[w] = i;
        pl->row_for_nnz[w] = i;
        pl->adj_values[w] = 0.0f;
        pl->diag_mask[w] = 1.0f;
        ++w;
        pl->crow[i + 1] = w;
    }
    // build_transposed_pattern: counting sort of the entries by column
    pl->trans_crow.assign(R + 1, 0);
    for (int k = 0; k < nnz; ++k) pl->trans_crow[pl->col[k] + 1]++;
    for (int i = 0; i < R; ++i) pl->trans_crow[i + 1] += pl->trans_crow[i];
    pl->trans_col.resize(nnz);
    pl->trans_to_orig.resize(nnz);
    {
        std::vector<int> cur(pl->trans_crow.begin(), pl->trans_crow.end() - 1);
        for (int k = 0; k < nnz; ++k) {
            int pos = cur[pl->col[k]]++;
            pl->trans_col[pos] = pl->row_for_nnz[k];
            pl->trans_to_orig[pos] = k;
        }
    }

    // levels: longest upstream path (upstream ids are smaller)
    pl->level.assign(R, 0);
    int maxlev = 0;
    for (int i = 0; i < R; ++i) {
        int lv = 0;
        for (int k = pl->crow[i]; k < pl->crow[i + 1]; ++k)
            if (pl->col[k] != i) lv = std::max(lv, pl->level[pl->col[k]] + 1);
        pl->level[i] = lv;
        maxlev = std::max(maxlev, lv);
    }
    pl->nlevels = maxlev + 1;
    // level-major order: counting sort by level, stable in reach id
    std::vector<int> lc(maxlev + 2, 0);
    for (int i = 0; i < R; ++i) lc[pl->level[i] + 1]++;
    for (int l = 0; l <= maxlev; ++l) lc[l + 1] += lc[l];
    pl->perm.resize(R);
    std::vector<int> inv(R);
    for (int i = 0; i < R; ++i) {
        int p = lc[pl->level[i]]++;
        pl->perm[p] = i;
        inv[i] = p;
    }
    const int C = (R + 63) / 64;
    pl->C = C;
    std::vector<int> up_ptr(R + 1, 0), dn_ptr(R + 1, 0), diag_k(R, 0);
    std::vector<int> up_pos, up_k, dn_pos, dn_k;
    std::vector<float> up_adj, dn_adj;
    up_pos.reserve(nnz_off);
    dn_pos.reserve(nnz_off);
    for (int p = 0; p < R; ++p) {
        const int i = pl->perm[p];
        for (int k = pl->crow[i]; k < pl->crow[i + 1]; ++k) {
            if (pl->col[k] == i) {
                diag_k[p] = k;
            } else {
                up_pos.push_back(inv[pl->col[k]]);
                up_adj.push_back(pl->adj_values[k]);
                up_k.push_back(k);
            }
        }
        up_ptr[p + 1] = static_cast<int>(up_pos.size());
        for (int k = pl->trans_crow[i]; k < pl->trans_crow[i + 1]; ++k) {
            const int r = pl->trans_col[k];
            if (r == i) continue;
            const int ko = pl->trans_to_orig[k];
            dn_pos.push_back(inv[r]);
            dn_adj.push_back(pl->adj_values[ko]);
            dn_k.push_back(ko);
        }
        dn_ptr[p + 1] = static_cast<int>(dn_pos.size());
    }
    // sub-levels inside each chunk and staging capacity
    std::vector<unsigned char> sub(R, 0), subr(R, 0);
    std::vector<int> nsub(C, 1), nsubr(C, 1);
    for (int c = 0; c < C; ++c) {
        const int p0 = c * 64, p1 = std::min(p0 + 64, R);
        if (up_ptr[p1] - up_ptr[p0] > kStageMax || dn_ptr[p1] - dn_ptr[p0] > kStageMax) {
            delete pl;
            return fail(DDRS_ERR_CAPACITY, "a 64-reach chunk has more than " +
                                               std::to_string(kStageMax) +
                                               " upstream or downstream links");
        }
        int mx = 0;
        for (int p = p0; p < p1; ++p) {
            int s = 0;
            for (int e = up_ptr[p]; e < up_ptr[p + 1]; ++e)
                if (up_pos[e] >= p0) s = std::max(s, sub[up_pos[e]] + 1);
            sub[p] = static_cast<unsigned char>(s);
            mx = std::max(mx, s);
        }
        nsub[c] = mx + 1;
        mx = 0;
        for (int p = p1 - 1; p >= p0; --p) {
            int s = 0;
            for (int e = dn_ptr[p]; e < dn_ptr[p + 1]; ++e)
                if (dn_pos[e] < p1) s = std::max(s, subr[dn_pos[e]] + 1);
            subr[p] = static_cast<unsigned char>(s);
            mx = std::max(mx, s);
        }
        nsubr[c] = mx + 1;
    }

    pl->h_up_ptr.swap(up_ptr);
    pl->h_up_pos.swap(up_pos);
    pl->h_up_adj.swap(up_adj);
    pl->h_up_k.swap(up_k);
    pl->h_dn_ptr.swap(dn_ptr);
    pl->h_dn_pos.swap(dn_pos);
    pl->h_dn_adj.swap(dn_adj);
    pl->h_dn_k.swap(dn_k);
    pl->h_diag_k.swap(diag_k);
    pl->h_sub.swap(sub);
    pl->h_subr.swap(subr);
    pl->h_nsub.swap(nsub);
    pl->h_nsubr.swap(nsubr);
    *out = pl;
    return DDRS_OK;
}

void ddrs_plan_destroy(ddrs_plan_t plan) {
    if (!plan) return;
    for (void* p : plan->dev) hipFree(p);
    delete plan;
}

ddrs_status ddrs_plan_info(ddrs_plan_t plan, int32_t* n, int32_t* nnz, int32_t* n_chunks,
                           int32_t* n_levels) {
    if (!plan) return fail(DDRS_ERR_INVALID, "plan is NULL");
    if (n) *n = plan->R;
    if (nnz) *nnz = plan->nnz;
    if (n_chunks) *n_chunks = plan->C;
    if (n_levels) *n_levels = plan->nlevels;
    return DDRS_OK;
}

ddrs_status ddrs_plan_get_csr(ddrs_plan_t plan, int32_t* crow, int32_t* col, int32_t* rfn,
                              float* adj, float* dmask, int32_t* tcrow, int32_t* tcol,
                              int32_t* t2o) {
    if (!plan) return fail(DDRS_ERR_INVALID, "plan is NULL");
    auto cp = [](auto* dst, const auto& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(crow, plan->crow);
    cp(col, plan->col);
    cp(rfn, plan->row_for_nnz);
    cp(adj, plan->adj_values);
    cp(dmask, plan->diag_mask);
    cp(tcrow, plan->trans_crow);
    cp(tcol, plan->trans_col);
    cp(t2o, plan->trans_to_orig);
    return DDRS_OK;
}

ddrs_status ddrs_plan_get_schedule(ddrs_plan_t plan, int32_t* h_perm, int32_t* h_level) {
    if (!plan) return fail(DDRS_ERR_INVALID, "plan is NULL");
    if (h_perm) std::memcpy(h_perm, plan->perm.data(), plan->perm.size() * sizeof(int));
    if (h_level) std::memcpy(h_level, plan->level.data(), plan->level.size() * sizeof(int));
    return DDRS_OK;
}

}  // extern "C"

namespace {

// Upload the schedule on first use by a workspace (plan creation itself is
// host-only work).
ddrs_status plan_to_device(ddrs_plan* pl) {
    if (pl->on_device) return DDRS_OK;
    Sched& sc = pl->sc;
    sc.R = pl->R;
    sc.C = pl->C;
    int *d_perm, *d_up_ptr, *d_up_pos, *d_up_k, *d_dn_ptr, *d_dn_pos, *d_dn_k, *d_diag, *d_nsub,
        *d_nsubr;
    float *d_up_adj, *d_dn_adj;
    unsigned char *d_sub, *d_subr;
    hipError_t e = hipSuccess;
#define UP(dp, hv)                               \
    if (e == hipSuccess) {                       \
        e = upload(&dp, hv);                     \
        if (dp) pl->dev.push_back(dp);           \
    }
    UP(d_perm, pl->perm);
    UP(d_up_ptr, pl->h_up_ptr);
    UP(d_up_pos, pl->h_up_pos);
    UP(d_up_adj, pl->h_up_adj);
    UP(d_up_k, pl->h_up_k);
    UP(d_dn_ptr, pl->h_dn_ptr);
    UP(d_dn_pos, pl->h_dn_pos);
    UP(d_dn_adj, pl->h_dn_adj);
    UP(d_dn_k, pl->h_dn_k);
    UP(d_diag, pl->h_diag_k);
    UP(d_sub, pl->h_sub);
    UP(d_subr, pl->h_subr);
    UP(d_nsub, pl->h_nsub);
    UP(d_nsubr, pl->h_nsubr);
    UP(pl->d_row_for_nnz, pl->row_for_nnz);
    UP(pl->d_col, pl->col);
#undef UP
    if (e != hipSuccess) return fail(DDRS_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e));
    sc.perm = d_perm;
    sc.up_ptr = d_up_ptr;
    sc.up_pos = d_up_pos;
    sc.up_adj = d_up_adj;
    sc.up_k = d_up_k;
    sc.dn_ptr = d_dn_ptr;
    sc.dn_pos = d_dn_pos;
    sc.dn_adj = d_dn_adj;
    sc.dn_k = d_dn_k;
    sc.diag_k = d_diag;
    sc.sub = d_sub;
    sc.subr = d_subr;
    sc.nsub = d_nsub;
    sc.nsubr = d_nsubr;
    pl->on_device = true;
    return DDRS_OK;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// workspace

ddrs_status ddrs_workspace_create(ddrs_plan_t plan, int32_t window, int32_t max_steps,
                                  ddrs_workspace_t* out) {
    if (!out) return fail(DDRS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!plan) return fail(DDRS_ERR_INVALID, "plan is NULL");
    if (window < 1 || max_steps < 0) return fail(DDRS_ERR_INVALID, "window must be >= 1");
    ddrs_status pst = plan_to_device(plan);
    if (pst != DDRS_OK) return pst;
    auto* ws = new ddrs_workspace();
    ws->plan = plan;
    ws->window = window;
    ws->max_steps = max_steps;
    ws->max_seg = std::max(1, (max_steps + window - 1) / window);
    const size_t R = plan->R;
    ddrs_status st = DDRS_OK;
    auto F = [&](float** p, size_t n) {
        if (st == DDRS_OK) st = dalloc(ws, reinterpret_cast<void**>(p), n * sizeof(float));
    };
    F(&ws->st.n, R);
    F(&ws->st.qsp, R);
    F(&ws->st.p, R);
    F(&ws->st.len, R);
    F(&ws->st.slope, R);
    F(&ws->st.sqrt_s, R);
    F(&ws->st.xst, R);
    F(&ws->st_q, R);
    F(&ws->st_i, R);
    F(&ws->st_lam, R);
    F(&ws->acc_n, R);
    F(&ws->acc_qsp, R);
    F(&ws->acc_p, R);
    F(&ws->sc_q, R);
    F(&ws->sc_i, R);
    F(&ws->sc_qp, R);
    F(&ws->ck, static_cast<size_t>(ws->max_seg) * 2 * R);
    if (st == DDRS_OK)
        st = dalloc(ws, reinterpret_cast<void**>(&ws->hist), static_cast<size_t>(window) * R * 8);
    if (st == DDRS_OK)
        st = dalloc(ws, reinterpret_cast<void**>(&ws->amsg), static_cast<size_t>(window) * R * 16);
    if (st == DDRS_OK) st = dalloc(ws, reinterpret_cast<void**>(&ws->slot), R * 8);
    if (st == DDRS_OK) st = dalloc(ws, reinterpret_cast<void**>(&ws->err), sizeof(int));
    if (st != DDRS_OK) {
        ddrs_workspace_destroy(ws);
        return st;
    }
    // tags start at epoch 0 = "never written"
    hipError_t e = hipMemset(ws->hist, 0, static_cast<size_t>(window) * R * 8);
    if (e == hipSuccess) e = hipMemset(ws->amsg, 0, static_cast<size_t>(window) * R * 16);
    if (e == hipSuccess) e = hipMemset(ws->slot, 0, R * 8);
    if (e == hipSuccess) e = hipMemset(ws->err, 0, sizeof(int));
    if (e != hipSuccess) {
        ddrs_workspace_destroy(ws);
        return fail(DDRS_ERR_HIP, std::string("workspace init: ") + hipGetErrorString(e));
    }
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    auto grid = [&](int which) -> int {
        int nb = occupancy_blocks(which);
        long cap = static_cast<long>(nb) * ncu;
        return static_cast<int>(std::min<long>(cap, plan->C));
    };
    for (int v = 0; v < 3; ++v) {
        ws->grid_fwd[v] = grid(v);
        ws->grid_adj[v] = grid(3 + v);
    }
    ws->grid_tri_fwd = grid(6);
    ws->grid_tri_adj = grid(7);
    int gmin = std::min({ws->grid_fwd[0], ws->grid_fwd[1], ws->grid_fwd[2], ws->grid_adj[0],
                         ws->grid_adj[1], ws->grid_adj[2], ws->grid_tri_fwd, ws->grid_tri_adj});
    if (gmin <= 0) {
        ddrs_workspace_destroy(ws);
        return fail(DDRS_ERR_HIP, "occupancy query returned no resident workgroups");
    }
    if ((plan->C + gmin - 1) / gmin > kMaxChunksPerWave) {
        ddrs_workspace_destroy(ws);
        return fail(DDRS_ERR_CAPACITY, "network too large for one GPU's resident workgroups");
    }
    *out = ws;
    return DDRS_OK;
}

void ddrs_workspace_destroy(ddrs_workspace_t ws) {
    if (!ws) return;
    for (void* p : ws->dev) hipFree(p);
    for (auto e : ws->ev_fwd) hipEventDestroy(e);
    for (auto e : ws->ev_adj) hipEventDestroy(e);
    delete ws;
}

ddrs_status ddrs_workspace_bytes(ddrs_workspace_t ws, size_t* bytes) {
    if (!ws || !bytes) return fail(DDRS_ERR_INVALID, "NULL argument");
    *bytes = ws->bytes;
    return DDRS_OK;
}

ddrs_status ddrs_workspace_check(ddrs_workspace_t ws, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    HIPCHK(hipStreamSynchronize(S(stream)));
    int flag = 0;
    HIPCHK(hipMemcpy(&flag, ws->err, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        hipMemset(ws->err, 0, sizeof(int));
        return fail(DDRS_ERR_HIP,
                    "persistent kernel watchdog: a workgroup made no progress for 2 s "
                    "(results of the last call are invalid)");
    }
    return DDRS_OK;
}

ddrs_status ddrs_mc_bind_statics(ddrs_workspace_t ws, const ddrs_params* params,
                                 const float* d_n, const float* d_q_spatial,
                                 const float* d_p_spatial, const float* d_length,
                                 const float* d_slope, const float* d_x_storage, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!d_n || !d_q_spatial || !d_p_spatial || !d_length || !d_slope)
        return fail(DDRS_ERR_INVALID, "statics must not be NULL (x_storage may be)");
    launch_bind_statics(ws->plan->R, ws->plan->sc.perm, d_n, d_q_spatial, d_p_spatial, d_length,
                        d_slope, d_x_storage, params->slope_lb, ws->st, S(stream));
    HIPCHK(hipGetLastError());
    ws->statics_bound = true;
    return DDRS_OK;
}

// ---------------------------------------------------------------------------
// single step

static FwdArgs fwd_args(ddrs_workspace* ws, const ddrs_params* p) {
    FwdArgs a{};
    a.sc = ws->plan->sc;
    a.st = ws->st;
    a.P = to_phys(p);
    a.hist = ws->hist;
    a.st_q = ws->st_q;
    a.st_i = ws->st_i;
    a.err = ws->err;
    return a;
}

static AdjArgs adj_args(ddrs_workspace* ws, const ddrs_params* p) {
    AdjArgs a{};
    a.sc = ws->plan->sc;
    a.st = ws->st;
    a.P = to_phys(p);
    a.hist = ws->hist;
    a.amsg = ws->amsg;
    a.st_lam = ws->st_lam;
    a.acc_n = ws->acc_n;
    a.acc_qsp = ws->acc_qsp;
    a.acc_p = ws->acc_p;
    a.sc_q = ws->sc_q;
    a.sc_i = ws->sc_i;
    a.sc_qp = ws->sc_qp;
    a.err = ws->err;
    return a;
}

ddrs_status ddrs_mc_step_fwd(ddrs_workspace_t ws, const ddrs_params* params, const float* d_q_t,
                             const float* d_q_prime_t, float* d_q_next, float* d_x_sol,
                             void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!ws->statics_bound) return fail(DDRS_ERR_INVALID, "statics not bound");
    if (!d_q_t || !d_q_prime_t || !d_q_next) return fail(DDRS_ERR_INVALID, "NULL input/output");
    hipStream_t s = S(stream);
    launch_state(ws->plan->sc, d_q_t, 0, params->discharge_lb, ws->st_q, ws->st_i, nullptr, s);
    FwdArgs a = fwd_args(ws, params);
    a.t0 = 0;
    a.s_base = 0;
    a.nsteps = 1;
    a.epoch = ++ws->epoch;
    a.qprime = d_q_prime_t;
    a.X_out = d_x_sol;
    a.Q_out = nullptr;
    launch_fwd(a, params, ws->grid_fwd[variant(params)], s);
    HIPCHK(hipGetLastError());
    launch_unpermute(ws->plan->R, ws->plan->sc.perm, ws->st_q, d_q_next, s);
    HIPCHK(hipGetLastError());
    ws->last_T = -1;
    return DDRS_OK;
}

ddrs_status ddrs_mc_step_bwd(ddrs_workspace_t ws, const ddrs_params* params, const float* d_q_t,
                             const float* d_q_prime_t, const float* d_x_sol,
                             const float* d_grad_q_next, float* d_grad_n,
                             float* d_grad_q_spatial, float* d_grad_p_spatial,
                             float* d_grad_q_t, float* d_grad_q_prime_t, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!ws->statics_bound) return fail(DDRS_ERR_INVALID, "statics not bound");
    if (!d_q_t || !d_q_prime_t || !d_x_sol || !d_grad_q_next)
        return fail(DDRS_ERR_INVALID, "NULL input");
    hipStream_t s = S(stream);
    const Sched& sc = ws->plan->sc;
    const int R = sc.R;
    launch_state(sc, d_q_t, 0, params->discharge_lb, ws->st_q, ws->st_i, nullptr, s);
    launch_hist_from_x(R, sc.perm, d_x_sol, ws->hist, s);
    launch_lam_init(R, sc.perm, d_grad_q_next, ws->st_lam, ws->acc_n, ws->acc_qsp, ws->acc_p, s);
    AdjArgs a = adj_args(ws, params);
    a.t0 = 0;
    a.s_base = 0;
    a.nsteps = 1;
    a.epoch = ++ws->epoch;
    a.qprime = d_q_prime_t;
    a.gQ = nullptr;
    a.gqp_out = d_grad_q_prime_t;  // row 0, c4 * gb without the q' clamp mask
    a.mask_qp = 0;
    a.ck_q = ws->st_q;
    a.ck_i = ws->st_i;
    launch_adj(a, params, ws->grid_adj[variant(params)], s);
    HIPCHK(hipGetLastError());
    if (d_grad_n) launch_unpermute(R, sc.perm, ws->acc_n, d_grad_n, s);
    if (d_grad_q_spatial) launch_unpermute(R, sc.perm, ws->acc_qsp, d_grad_q_spatial, s);
    if (d_grad_p_spatial) launch_unpermute(R, sc.perm, ws->acc_p, d_grad_p_spatial, s);
    if (d_grad_q_t) launch_unpermute(R, sc.perm, ws->st_lam, d_grad_q_t, s);
    HIPCHK(hipGetLastError());
    ws->last_T = -1;
    return DDRS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// windows

namespace {

// Forward of checkpoint segment `seg` (steps [seg*window, ...)) from the state
// currently in st_q / st_i; writes the segment's solve history.
ddrs_status run_fwd_segment(ddrs_workspace* ws, const ddrs_params* params, int T, int seg,
                            const float* qprime, float* Q_out, ddrs_form form, hipStream_t s,
                            bool timed) {
    const int steps = T - 1;
    const int t0 = seg * ws->window;
    const int len = std::min(ws->window, steps - t0);
    FwdArgs a = fwd_args(ws, params);
    a.t0 = t0;
    a.qprime = qprime;
    a.Q_out = Q_out;
    const int grid = ws->grid_fwd[variant(params)];
    if (form == DDRS_FORM_PERSISTENT) {
        a.s_base = 0;
        a.nsteps = len;
        a.epoch = ++ws->epoch;
        if (timed) record(ws, ws->ev_fwd, ws->n_fwd, s);
        launch_fwd(a, params, grid, s);
        if (timed) record(ws, ws->ev_fwd, ws->n_fwd, s);
    } else {
        for (int k = 0; k < len; ++k) {
            a.s_base = k;
            a.nsteps = 1;
            a.epoch = ++ws->epoch;
            if (timed) record(ws, ws->ev_fwd, ws->n_fwd, s);
            launch_fwd(a, params, grid, s);
            if (timed) record(ws, ws->ev_fwd, ws->n_fwd, s);
        }
    }
    HIPCHK(hipGetLastError());
    ws->hist_seg = seg;
    return DDRS_OK;
}

float* ck_q(ddrs_workspace* ws, int seg) { return ws->ck + static_cast<size_t>(seg) * 2 * ws->plan->R; }
float* ck_i(ddrs_workspace* ws, int seg) { return ck_q(ws, seg) + ws->plan->R; }

}  // namespace

extern "C" {

ddrs_status ddrs_mc_window_fwd(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                               const float* d_q_prime, const float* d_q0, float* d_Q_out,
                               ddrs_form form, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!ws->statics_bound) return fail(DDRS_ERR_INVALID, "statics not bound");
    if (T < 1) return fail(DDRS_ERR_INVALID, "T must be >= 1");
    if (T - 1 > ws->max_steps)
        return fail(DDRS_ERR_CAPACITY, "T-1 exceeds the workspace's max_steps");
    if (!d_q0 || (T > 1 && !d_q_prime)) return fail(DDRS_ERR_INVALID, "NULL input");
    if (form != DDRS_FORM_PLAIN && form != DDRS_FORM_PERSISTENT)
        return fail(DDRS_ERR_INVALID, "unknown form");
    hipStream_t s = S(stream);
    const int R = ws->plan->R;
    ws->n_fwd = 0;
    launch_state(ws->plan->sc, d_q0, 1, params->discharge_lb, ws->st_q, ws->st_i, d_Q_out, s);
    const int steps = T - 1;
    const int nseg = (steps + ws->window - 1) / ws->window;
    for (int j = 0; j < nseg; ++j) {
        HIPCHK(hipMemcpyAsync(ck_q(ws, j), ws->st_q, R * sizeof(float), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ck_i(ws, j), ws->st_i, R * sizeof(float), hipMemcpyDeviceToDevice, s));
        st = run_fwd_segment(ws, params, T, j, d_q_prime, d_Q_out, form, s, true);
        if (st != DDRS_OK) return st;
    }
    ws->last_T = T;
    if (nseg == 0) ws->hist_seg = -1;
    return DDRS_OK;
}

ddrs_status ddrs_mc_window_adj(ddrs_workspace_t ws, const ddrs_params* params, int32_t T,
                               const float* d_q_prime, const float* d_q0,
                               const float* d_grad_Q, float* d_grad_n,
                               float* d_grad_q_spatial, float* d_grad_p_spatial,
                               float* d_grad_q_prime, float* d_grad_q0, ddrs_form form,
                               void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (ws->last_T != T) return fail(DDRS_ERR_INVALID, "no matching ddrs_mc_window_fwd before the adjoint");
    if (!d_q0 || (T > 1 && !d_q_prime)) return fail(DDRS_ERR_INVALID, "NULL input");
    if (form != DDRS_FORM_PLAIN && form != DDRS_FORM_PERSISTENT)
        return fail(DDRS_ERR_INVALID, "unknown form");
    hipStream_t s = S(stream);
    const Sched& sc = ws->plan->sc;
    const int R = sc.R;
    ws->n_adj = 0;
    const float* g_last = d_grad_Q ? d_grad_Q + static_cast<size_t>(T - 1) * R : nullptr;
    launch_lam_init(R, sc.perm, g_last, ws->st_lam, ws->acc_n, ws->acc_qsp, ws->acc_p, s);
    if (d_grad_q_prime) launch_zero(static_cast<size_t>(R), d_grad_q_prime + static_cast<size_t>(T - 1) * R, s);
    const int steps = T - 1;
    const int nseg = (steps + ws->window - 1) / ws->window;
    const int grid = ws->grid_adj[variant(params)];
    for (int j = nseg - 1; j >= 0; --j) {
        if (ws->hist_seg != j) {
            HIPCHK(hipMemcpyAsync(ws->st_q, ck_q(ws, j), R * sizeof(float), hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(ws->st_i, ck_i(ws, j), R * sizeof(float), hipMemcpyDeviceToDevice, s));
            st = run_fwd_segment(ws, params, T, j, d_q_prime, nullptr, DDRS_FORM_PERSISTENT, s, false);
            if (st != DDRS_OK) return st;
        }
        const int t0 = j * ws->window;
        const int len = std::min(ws->window, steps - t0);
        AdjArgs a = adj_args(ws, params);
        a.t0 = t0;
        a.qprime = d_q_prime;
        a.gQ = d_grad_Q;
        a.gqp_out = d_grad_q_prime;
        a.mask_qp = 1;
        a.ck_q = ck_q(ws, j);
        a.ck_i = ck_i(ws, j);
        if (form == DDRS_FORM_PERSISTENT) {
            a.s_base = 0;
            a.nsteps = len;
            a.epoch = ++ws->epoch;
            record(ws, ws->ev_adj, ws->n_adj, s);
            launch_adj(a, params, grid, s);
            record(ws, ws->ev_adj, ws->n_adj, s);
        } else {
            for (int k = len - 1; k >= 0; --k) {
                a.s_base = k;
                a.nsteps = 1;
                a.epoch = ++ws->epoch;
                record(ws, ws->ev_adj, ws->n_adj, s);
                launch_adj(a, params, grid, s);
                record(ws, ws->ev_adj, ws->n_adj, s);
            }
        }
        HIPCHK(hipGetLastError());
    }
    if (d_grad_q0) launch_grad_q0(R, sc.perm, ws->st_lam, d_q0, params->discharge_lb, d_grad_q0, s);
    if (d_grad_n) launch_unpermute(R, sc.perm, ws->acc_n, d_grad_n, s);
    if (d_grad_q_spatial) launch_unpermute(R, sc.perm, ws->acc_qsp, d_grad_q_spatial, s);
    if (d_grad_p_spatial) launch_unpermute(R, sc.perm, ws->acc_p, d_grad_p_spatial, s);
    HIPCHK(hipGetLastError());
    return DDRS_OK;
}

// ---------------------------------------------------------------------------
// sparse seam

ddrs_status ddrs_tri_solve_fwd(ddrs_workspace_t ws, const float* d_a_values, const float* d_b,
                               float* d_x, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    if (!d_a_values || !d_b || !d_x) return fail(DDRS_ERR_INVALID, "NULL argument");
    TriArgs a{};
    a.sc = ws->plan->sc;
    a.epoch = ++ws->epoch;
    a.a_values = d_a_values;
    a.b = d_b;
    a.x = d_x;
    a.slot = ws->slot;
    a.err = ws->err;
    launch_tri_fwd(a, ws->grid_tri_fwd, S(stream));
    HIPCHK(hipGetLastError());
    return DDRS_OK;
}

ddrs_status ddrs_tri_solve_adj(ddrs_workspace_t ws, const float* d_a_values, const float* d_x,
                               const float* d_grad_x, float* d_grad_b, float* d_grad_a,
                               void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    if (!d_a_values || !d_grad_x || !d_grad_b || (d_grad_a && !d_x))
        return fail(DDRS_ERR_INVALID, "NULL argument");
    hipStream_t s = S(stream);
    TriArgs a{};
    a.sc = ws->plan->sc;
    a.epoch = ++ws->epoch;
    a.a_values = d_a_values;
    a.b = d_grad_x;
    a.x = d_grad_b;
    a.slot = ws->slot;
    a.err = ws->err;
    launch_tri_adj(a, ws->grid_tri_adj, s);
    HIPCHK(hipGetLastError());
    if (d_grad_a) {
        launch_grad_a(ws->plan->nnz, ws->plan->d_row_for_nnz, ws->plan->d_col, d_grad_b, d_x,
                      d_grad_a, s);
        HIPCHK(hipGetLastError());
    }
    return DDRS_OK;
}

ddrs_status ddrs_hotstart(ddrs_workspace_t ws, const ddrs_params* params,
                          const float* d_q_prime0, float* d_q0, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!d_q_prime0 || !d_q0) return fail(DDRS_ERR_INVALID, "NULL argument");
    TriArgs a{};
    a.sc = ws->plan->sc;
    a.epoch = ++ws->epoch;
    a.a_values = nullptr;  // A = I - N
    a.b = d_q_prime0;
    a.x = d_q0;
    a.slot = ws->slot;
    a.clamp = 1;
    a.lb = params->discharge_lb;
    a.err = ws->err;
    launch_tri_fwd(a, ws->grid_tri_fwd, S(stream));
    HIPCHK(hipGetLastError());
    return DDRS_OK;
}

ddrs_status ddrs_hotstart_adj(ddrs_workspace_t ws, const ddrs_params* params, const float* d_q0,
                              const float* d_grad_q0, float* d_grad_q_prime0, void* stream) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ddrs_status st = check_params(params);
    if (st != DDRS_OK) return st;
    if (!d_q0 || !d_grad_q0 || !d_grad_q_prime0) return fail(DDRS_ERR_INVALID, "NULL argument");
    TriArgs a{};
    a.sc = ws->plan->sc;
    a.epoch = ++ws->epoch;
    a.a_values = nullptr;
    a.b = d_grad_q0;
    a.mask_x = d_q0;
    a.lb = params->discharge_lb;
    a.x = d_grad_q_prime0;
    a.slot = ws->slot;
    a.err = ws->err;
    launch_tri_adj(a, ws->grid_tri_adj, S(stream));
    HIPCHK(hipGetLastError());
    return DDRS_OK;
}

// ---------------------------------------------------------------------------
// timing

ddrs_status ddrs_set_timing(ddrs_workspace_t ws, int32_t enable) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    ws->timing = enable != 0;
    ws->n_fwd = ws->n_adj = 0;
    return DDRS_OK;
}

ddrs_status ddrs_get_timing(ddrs_workspace_t ws, int32_t* fwd_launches, float* fwd_ms,
                            int32_t* adj_launches, float* adj_ms) {
    if (!ws) return fail(DDRS_ERR_INVALID, "workspace is NULL");
    auto sum = [&](std::vector<hipEvent_t>& pool, int n, float* ms) -> ddrs_status {
        float tot = 0.0f;
        if (n > 0) HIPCHK(hipEventSynchronize(pool[n - 1]));
        for (int k = 0; k + 1 < n; k += 2) {
            float v = 0.0f;
            HIPCHK(hipEventElapsedTime(&v, pool[k], pool[k + 1]));
            tot += v;
        }
        if (ms) *ms = tot;
        return DDRS_OK;
    };
    ddrs_status st = sum(ws->ev_fwd, ws->n_fwd, fwd_ms);
    if (st != DDRS_OK) return st;
    st = sum(ws->ev_adj, ws->n_adj, adj_ms);
    if (st != DDRS_OK) return st;
    if (fwd_launches) *fwd_launches = ws->n_fwd / 2;
    if (adj_launches) *adj_launches = ws->n_adj / 2;
    return DDRS_OK;
}

}  // extern "C"

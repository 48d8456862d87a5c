// Host index construction (bit-exact with the reference):
//   ddrs_build_topology  adjacency/build.rs:170-272 (topological order of a
//                        MERIT-style fabric, rustworkx/petgraph DFS finish-time
//                        rule, cycle drop and retry, isolated COMIDs appended)
//   ddrs_compress        data/collate.rs:197-349 (sorted active set, compressed
//                        lower-triangular COO, gauge positions)
//   ddrs_compress_ex     the same with the compressed reach-subdivision map
//                        (collate.rs:141-180) and outflow_idx (297-336)
#include <algorithm>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/ddrs_hip.h"

extern "C" void ddrs_internal_set_error(const char* msg);

namespace {

ddrs_status ifail(ddrs_status s, const std::string& m) {
    ddrs_internal_set_error(m.c_str());
    return s;
}

struct Rec {
    int64_t comid;
    std::vector<int64_t> ups;
};

// Graph with node insertion order = sorted downstream keys, each followed by
// its sorted upstream COMIDs (build.rs:307-348).
struct Graph {
    std::vector<int64_t> node;
    std::vector<std::vector<int>> succ;

    explicit Graph(const std::vector<Rec>& recs) {
        std::unordered_map<int64_t, std::vector<int64_t>> ups;
        for (const Rec& r : recs)
            for (int64_t u : r.ups)
                if (u > 0) ups[r.comid].push_back(u);
        std::vector<int64_t> keys;
        keys.reserve(ups.size());
        for (auto& kv : ups) {
            std::sort(kv.second.begin(), kv.second.end());
            kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
            keys.push_back(kv.first);
        }
        std::sort(keys.begin(), keys.end());
        std::unordered_map<int64_t, int> idx;
        auto add = [&](int64_t c) {
            if (idx.emplace(c, static_cast<int>(node.size())).second) node.push_back(c);
        };
        for (int64_t k : keys) {
            add(k);
            for (int64_t u : ups[k]) add(u);
        }
        succ.assign(node.size(), {});
        for (int64_t k : keys) {
            const int ti = idx[k];
            for (int64_t u : ups[k]) succ[idx[u]].push_back(ti);
        }
    }

    // COMIDs on simple cycles: members of SCCs larger than one, plus
    // self-loops (Tarjan, iterative).
    std::vector<int64_t> cycle_comids() const {
        const int n = static_cast<int>(node.size());
        std::vector<int> index(n, -1), low(n, 0);
        std::vector<char> on(n, 0), cyc(n, 0);
        std::vector<int> stack;
        int counter = 0;
        for (int s0 = 0; s0 < n; ++s0) {
            if (index[s0] >= 0) continue;
            std::vector<std::pair<int, int>> work{{s0, 0}};
            while (!work.empty()) {
                int v = work.back().first;
                int ci = work.back().second;
                if (ci == 0 && index[v] < 0) {
                    index[v] = low[v] = counter++;
                    stack.push_back(v);
                    on[v] = 1;
                }
                if (ci < static_cast<int>(succ[v].size())) {
                    const int w = succ[v][ci];
                    work.back().second++;
                    if (index[w] < 0) work.push_back({w, 0});
                    else if (on[w]) low[v] = std::min(low[v], index[w]);
                } else {
                    if (low[v] == index[v]) {
                        std::vector<int> comp;
                        while (true) {
                            int w = stack.back();
                            stack.pop_back();
                            on[w] = 0;
                            comp.push_back(w);
                            if (w == v) break;
                        }
                        if (comp.size() > 1)
                            for (int w : comp) cyc[w] = 1;
                    }
                    work.pop_back();
                    if (!work.empty()) {
                        const int p = work.back().first;
                        low[p] = std::min(low[p], low[v]);
                    }
                }
            }
        }
        for (int i = 0; i < n; ++i)
            for (int s : succ[i])
                if (s == i) cyc[i] = 1;
        std::vector<int64_t> out;
        for (int i = 0; i < n; ++i)
            if (cyc[i]) out.push_back(node[i]);
        std::sort(out.begin(), out.end());
        return out;
    }

    // Iterative DFS: outer loop in node-index order; on first visit push
    // undiscovered successors in reverse insertion order; record a node when it
    // is popped the second time; reverse the finish stack (build.rs:379-427).
    std::vector<int> toposort() const {
        const int n = static_cast<int>(node.size());
        std::vector<char> disc(n, 0), fin(n, 0);
        std::vector<int> finish, stack;
        finish.reserve(n);
        for (int s0 = 0; s0 < n; ++s0) {
            if (disc[s0]) continue;
            stack.push_back(s0);
            while (!stack.empty()) {
                const int v = stack.back();
                if (!disc[v]) {
                    disc[v] = 1;
                    for (auto it = succ[v].rbegin(); it != succ[v].rend(); ++it)
                        if (!disc[*it]) stack.push_back(*it);
                } else {
                    stack.pop_back();
                    if (!fin[v]) {
                        fin[v] = 1;
                        finish.push_back(v);
                    }
                }
            }
        }
        std::reverse(finish.begin(), finish.end());
        return finish;
    }
};

ddrs_status build(const std::vector<Rec>& recs, const std::vector<int64_t>& all,
                  std::vector<int64_t>& order, std::vector<int>& rows, std::vector<int>& cols,
                  std::vector<int64_t>& dropped) {
    Graph g(recs);
    std::vector<int64_t> cyc = g.cycle_comids();
    if (!cyc.empty()) {
        std::unordered_set<int64_t> drop(cyc.begin(), cyc.end());
        std::vector<Rec> filt;
        std::vector<int64_t> fall;
        for (const Rec& r : recs)
            if (!drop.count(r.comid)) {
                filt.push_back(r);
                fall.push_back(r.comid);
            }
        ddrs_status st = build(filt, fall, order, rows, cols, dropped);
        if (st != DDRS_OK) return st;
        for (int64_t c : cyc) drop.insert(c);
        for (int64_t c : dropped) drop.insert(c);
        dropped.assign(drop.begin(), drop.end());
        std::sort(dropped.begin(), dropped.end());
        return DDRS_OK;
    }
    std::vector<int> ts = g.toposort();
    order.clear();
    for (int i : ts) order.push_back(g.node[i]);
    std::unordered_set<int64_t> connected(order.begin(), order.end());
    std::vector<int64_t> iso;
    for (int64_t c : all)
        if (!connected.count(c)) iso.push_back(c);
    std::sort(iso.begin(), iso.end());
    iso.erase(std::unique(iso.begin(), iso.end()), iso.end());
    order.insert(order.end(), iso.begin(), iso.end());
    std::unordered_map<int64_t, int> pos;
    for (size_t k = 0; k < order.size(); ++k) pos[order[k]] = static_cast<int>(k);
    rows.clear();
    cols.clear();
    for (int v : ts) {
        if (g.succ[v].empty()) continue;
        if (g.succ[v].size() != 1)
            return ifail(DDRS_ERR_INVALID, "node COMID " + std::to_string(g.node[v]) + " has " +
                                               std::to_string(g.succ[v].size()) +
                                               " successors, not dendritic");
        cols.push_back(pos[g.node[v]]);
        rows.push_back(pos[g.node[g.succ[v][0]]]);
    }
    for (size_t k = 0; k < rows.size(); ++k)
        if (rows[k] < cols[k]) return ifail(DDRS_ERR_INVALID, "COO is not lower triangular");
    dropped.clear();
    return DDRS_OK;
}

}  // namespace

extern "C" {

ddrs_status ddrs_build_topology(int32_t n_records, const int64_t* h_records, int64_t* h_order,
                                int32_t order_cap, int32_t* n_order, int32_t* h_rows,
                                int32_t* h_cols, int32_t edge_cap, int32_t* n_edges,
                                int64_t* h_dropped, int32_t dropped_cap, int32_t* n_dropped) {
    if (n_records < 0 || (n_records > 0 && !h_records))
        return ifail(DDRS_ERR_INVALID, "bad records");
    std::vector<Rec> recs(n_records);
    std::vector<int64_t> all(n_records);
    for (int i = 0; i < n_records; ++i) {
        recs[i].comid = h_records[5 * i];
        for (int k = 1; k < 5; ++k)
            if (h_records[5 * i + k] > 0) recs[i].ups.push_back(h_records[5 * i + k]);
        all[i] = recs[i].comid;
    }
    std::vector<int64_t> order, dropped;
    std::vector<int> rows, cols;
    ddrs_status st = build(recs, all, order, rows, cols, dropped);
    if (st != DDRS_OK) return st;
    if (static_cast<int64_t>(order.size()) > order_cap ||
        static_cast<int64_t>(rows.size()) > edge_cap ||
        static_cast<int64_t>(dropped.size()) > dropped_cap)
        return ifail(DDRS_ERR_CAPACITY, "output capacity too small");
    std::copy(order.begin(), order.end(), h_order);
    std::copy(rows.begin(), rows.end(), h_rows);
    std::copy(cols.begin(), cols.end(), h_cols);
    std::copy(dropped.begin(), dropped.end(), h_dropped);
    if (n_order) *n_order = static_cast<int32_t>(order.size());
    if (n_edges) *n_edges = static_cast<int32_t>(rows.size());
    if (n_dropped) *n_dropped = static_cast<int32_t>(dropped.size());
    return DDRS_OK;
}

ddrs_status ddrs_compress(int32_t n_edges_in, const int32_t* h_rows_in, const int32_t* h_cols_in,
                          int32_t n_gauges, const int32_t* h_gauges, int32_t* h_active,
                          int32_t active_cap, int32_t* n_active, int32_t* h_rows,
                          int32_t* h_cols, int32_t* n_edges, int32_t* h_gauge) {
    if (n_edges_in < 0 || n_gauges < 0) return ifail(DDRS_ERR_INVALID, "negative sizes");
    std::vector<std::pair<int, int>> e;
    e.reserve(n_edges_in);
    for (int k = 0; k < n_edges_in; ++k) e.emplace_back(h_rows_in[k], h_cols_in[k]);
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    std::vector<int> act;
    act.reserve(2 * e.size() + n_gauges);
    for (auto& rc : e) {
        act.push_back(rc.first);
        act.push_back(rc.second);
    }
    for (int g = 0; g < n_gauges; ++g) act.push_back(h_gauges[g]);
    std::sort(act.begin(), act.end());
    act.erase(std::unique(act.begin(), act.end()), act.end());
    if (act.empty()) return ifail(DDRS_ERR_INVALID, "compress: empty active set");
    if (static_cast<int64_t>(act.size()) > active_cap)
        return ifail(DDRS_ERR_CAPACITY, "active capacity too small");
    auto find = [&](int v) {
        return static_cast<int>(std::lower_bound(act.begin(), act.end(), v) - act.begin());
    };
    for (size_t k = 0; k < e.size(); ++k) {
        const int r = find(e[k].first), c = find(e[k].second);
        if (r < c) return ifail(DDRS_ERR_INVALID, "lower-triangular violated");
        h_rows[k] = r;
        h_cols[k] = c;
    }
    for (int g = 0; g < n_gauges; ++g) h_gauge[g] = find(h_gauges[g]);
    std::copy(act.begin(), act.end(), h_active);
    if (n_active) *n_active = static_cast<int32_t>(act.size());
    if (n_edges) *n_edges = static_cast<int32_t>(e.size());
    return DDRS_OK;
}

namespace {

// Owner of sub-reach row `row` in a strictly increasing parent map: the last
// offset <= row (collate.rs:124-129).
int parent_of_row(const int32_t* off, int n_off, int row) {
    return static_cast<int>(std::upper_bound(off, off + n_off, row) - off) - 1;
}

}  // namespace

ddrs_status ddrs_compress_ex(int32_t n_edges_in, const int32_t* h_rows_in,
                             const int32_t* h_cols_in, int32_t n_gauges,
                             const int32_t* h_gauges, int32_t ddr_match,
                             int32_t n_conus_parent, const int32_t* h_conus_parent_offset,
                             int32_t* h_active, int32_t active_cap, int32_t* n_active,
                             int32_t* h_rows, int32_t* h_cols, int32_t* n_edges,
                             int32_t* h_gauge, int32_t* h_parent_offset, int32_t* n_parent,
                             int32_t* h_outflow_ptr, int32_t* h_outflow, int32_t outflow_cap,
                             int32_t* n_outflow) {
    if (n_conus_parent < 0 || (n_conus_parent > 0 && !h_conus_parent_offset))
        return ifail(DDRS_ERR_INVALID, "compress: bad parent map");
    if (n_gauges > 0 && !h_outflow_ptr) return ifail(DDRS_ERR_INVALID, "compress: NULL outflow_ptr");
    int32_t na = 0, ne = 0;
    ddrs_status st = ddrs_compress(n_edges_in, h_rows_in, h_cols_in, n_gauges, h_gauges, h_active,
                                   active_cap, &na, h_rows, h_cols, &ne, h_gauge);
    if (st != DDRS_OK) return st;
    if (n_active) *n_active = na;
    if (n_edges) *n_edges = ne;
    // 4b. the reach-subdivision map in compressed rows: every parent present
    // must be present whole, as a contiguous run of its CONUS rows
    // (collate.rs:141-180)
    std::vector<int32_t> poff;
    if (n_conus_parent > 0) {
        const int32_t* off = h_conus_parent_offset;
        const int n_off = n_conus_parent + 1;
        for (int k = 0; k + 1 < n_off; ++k)
            if (!(off[k + 1] > off[k]) || off[0] != 0)
                return ifail(DDRS_ERR_INVALID, "compress: parent map not strictly increasing from 0");
        const int n_rows = off[n_conus_parent];
        poff.push_back(0);
        int i = 0;
        while (i < na) {
            const int row = h_active[i];
            if (row >= n_rows)
                return ifail(DDRS_ERR_INVALID,
                             "compress: CONUS row " + std::to_string(row) +
                                 " is outside the parent map (which covers " +
                                 std::to_string(n_rows) + " rows)");
            const int p = parent_of_row(off, n_off, row);
            const int lo = off[p], hi = off[p + 1], m = hi - lo;
            bool whole = na - i >= m;
            for (int k = 0; whole && k < m; ++k) whole = h_active[i + k] == lo + k;
            if (!whole)
                return ifail(DDRS_ERR_INVALID,
                             "compress: parent " + std::to_string(p) + " owns CONUS rows " +
                                 std::to_string(lo) + ".." + std::to_string(hi) +
                                 ", but the active set does not hold them contiguously from "
                                 "compressed row " + std::to_string(i));
            i += m;
            poff.push_back(i);
        }
        if (static_cast<int64_t>(poff.size()) > static_cast<int64_t>(active_cap) + 1 ||
            !h_parent_offset)
            return ifail(DDRS_ERR_CAPACITY, "compress: parent map capacity too small");
        std::copy(poff.begin(), poff.end(), h_parent_offset);
    }
    if (n_parent) *n_parent = static_cast<int32_t>(poff.empty() ? 0 : poff.size() - 1);
    // 5. outflow_idx (collate.rs:297-336): ddr_match, the gauge's upstream
    // columns in edge order (its own reach when it has none); otherwise the
    // gauge's own reach, at its parent's outlet piece when there is a map
    std::vector<int32_t> flat;
    if (n_gauges > 0) h_outflow_ptr[0] = 0;
    for (int g = 0; g < n_gauges; ++g) {
        const int gc = h_gauge[g];
        const size_t before = flat.size();
        if (ddr_match) {
            // edges are sorted by (row, col): the gauge's row is one run
            const int32_t* r0 = std::lower_bound(h_rows, h_rows + ne, gc);
            for (const int32_t* r = r0; r < h_rows + ne && *r == gc; ++r)
                flat.push_back(h_cols[r - h_rows]);
            if (flat.size() == before) flat.push_back(gc);
        } else if (!poff.empty()) {
            const int p = parent_of_row(poff.data(), static_cast<int>(poff.size()), gc);
            flat.push_back(poff[p + 1] - 1);
        } else {
            flat.push_back(gc);
        }
        h_outflow_ptr[g + 1] = static_cast<int32_t>(flat.size());
    }
    if (static_cast<int64_t>(flat.size()) > outflow_cap || (!flat.empty() && !h_outflow))
        return ifail(DDRS_ERR_CAPACITY, "compress: outflow capacity too small");
    std::copy(flat.begin(), flat.end(), h_outflow);
    if (n_outflow) *n_outflow = static_cast<int32_t>(flat.size());
    return DDRS_OK;
}

}  // extern "C"

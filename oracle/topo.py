"""Index-construction oracle — test infrastructure only.

Pure-Python restatement (small inputs) of the reference's integer/index work on
the path, used only by tests/ as the checker of the product's native builders:

* build_topo_and_coo  — adjacency/build.rs:170-272 (build_upstream_dict 275-292,
  DiGraph::build 307-348, topological_sort 379-427, cycle_comids 430-449,
  strongly_connected_components 452-501)
* compress            — data/collate.rs:197-349 (outflow_idx, ddr_match false
  without subdivision and ddr_match true)

Pinned by tests/golden/toposort_fuzz.jsonl (1000 graphs whose order was
produced by rustworkx 0.17.1 for the reference's own test,
tests/adjacency_build.rs:191-232).
"""
from __future__ import annotations


def build_upstream_dict(records):
    """records: iterable of (comid, [ups...]). downstream -> sorted unique ups > 0."""
    d: dict[int, list[int]] = {}
    for comid, ups in records:
        for up in ups:
            if up > 0:
                d.setdefault(comid, []).append(up)
    return {k: sorted(set(v)) for k, v in d.items()}


class DiGraph:
    def __init__(self, upstream_dict):
        node_index: dict[int, int] = {}
        self.node_data: list[int] = []

        def add(c):
            if c not in node_index:
                node_index[c] = len(self.node_data)
                self.node_data.append(c)

        keys = sorted(upstream_dict)
        for to in keys:
            add(to)
            for frm in upstream_dict[to]:
                add(frm)
        self.succ: list[list[int]] = [[] for _ in self.node_data]
        for to in keys:
            ti = node_index[to]
            for frm in upstream_dict[to]:
                self.succ[node_index[frm]].append(ti)

    def sccs(self):
        n = len(self.node_data)
        counter = 0
        index = [-1] * n
        low = [0] * n
        on = [False] * n
        stack: list[int] = []
        out = []
        for start in range(n):
            if index[start] != -1:
                continue
            work = [[start, 0]]
            while work:
                v, ci = work[-1]
                if ci == 0 and index[v] == -1:
                    index[v] = low[v] = counter
                    counter += 1
                    stack.append(v)
                    on[v] = True
                if ci < len(self.succ[v]):
                    w = self.succ[v][ci]
                    work[-1][1] += 1
                    if index[w] == -1:
                        work.append([w, 0])
                    elif on[w]:
                        low[v] = min(low[v], index[w])
                else:
                    if low[v] == index[v]:
                        comp = []
                        while True:
                            w = stack.pop()
                            on[w] = False
                            comp.append(w)
                            if w == v:
                                break
                        out.append(comp)
                    work.pop()
                    if work:
                        p = work[-1][0]
                        low[p] = min(low[p], low[v])
        return out

    def cycle_comids(self):
        on = set()
        for comp in self.sccs():
            if len(comp) > 1:
                on.update(comp)
        for i, s in enumerate(self.succ):
            if i in s:
                on.add(i)
        return sorted(self.node_data[i] for i in on)

    def toposort(self):
        cyc = self.cycle_comids()
        if cyc:
            return None, cyc
        n = len(self.node_data)
        disc = [False] * n
        fin = [False] * n
        finish = []
        for start in range(n):
            if disc[start]:
                continue
            stack = [start]
            while stack:
                nx = stack[-1]
                if not disc[nx]:
                    disc[nx] = True
                    for s in reversed(self.succ[nx]):
                        if not disc[s]:
                            stack.append(s)
                else:
                    stack.pop()
                    if not fin[nx]:
                        fin[nx] = True
                        finish.append(nx)
        finish.reverse()
        return finish, None


def build_topo_and_coo(records, all_comids=None):
    """Returns (order, rows, cols, dropped). records: list of (comid, [ups])."""
    records = list(records)
    if all_comids is None:
        all_comids = [c for c, _ in records]
    g = DiGraph(build_upstream_dict(records))
    ts, cyc = g.toposort()
    if ts is None:
        drop = set(cyc)
        filt = [(c, u) for c, u in records if c not in drop]
        order, rows, cols, more = build_topo_and_coo(filt, [c for c, _ in filt])
        drop.update(more)
        return order, rows, cols, sorted(drop)
    order = [g.node_data[i] for i in ts]
    connected = set(order)
    order += sorted({c for c in all_comids if c not in connected})
    idx = {c: k for k, c in enumerate(order)}
    rows, cols = [], []
    for node in ts:
        s = g.succ[node]
        if not s:
            continue
        if len(s) != 1:
            raise ValueError(f"node COMID {g.node_data[node]} has {len(s)} successors")
        cols.append(idx[g.node_data[node]])
        rows.append(idx[g.node_data[s[0]]])
    for r, c in zip(rows, cols):
        if r < c:
            raise ValueError("COO is not lower triangular")
    return order, rows, cols, []


def union_subgraphs(staids, subgraphs):
    """union_subgraphs (data/collate.rs:31-48). subgraphs: {staid: (gage_idx,
    gage_catchment, [(row, col), ...])}; a missing STAID is skipped. Returns
    (edges sorted and deduplicated, [(staid, gage_idx, gage_catchment)])."""
    edges, gauges = set(), []
    for s in staids:
        if s not in subgraphs:
            continue
        gi, gc, e = subgraphs[s]
        gauges.append((s, gi, gc))
        edges.update((int(r), int(c)) for r, c in e)
    return sorted(edges), gauges


def _parent_of_row(off, row):
    """The last offset <= row (collate.rs:124-129)."""
    p = 0
    while p + 1 < len(off) and off[p + 1] <= row:
        p += 1
    return p


def compressed_parent_offset(active, conus_off):
    """collate.rs:141-180: the CONUS parent map renumbered onto the active set;
    every present parent must be present whole and contiguous."""
    n_rows = conus_off[-1] if conus_off else 0
    out, i = [0], 0
    while i < len(active):
        row = active[i]
        if row >= n_rows:
            raise ValueError(f"compress: CONUS row {row} is outside the parent map")
        p = _parent_of_row(conus_off, row)
        lo, hi = conus_off[p], conus_off[p + 1]
        m = hi - lo
        if len(active) - i < m or any(active[i + k] != lo + k for k in range(m)):
            raise ValueError(f"compress: parent {p} is not held contiguously")
        i += m
        out.append(i)
    return out


def compress(edges, gauges, ddr_match=False, parent_offset=None):
    """collate::compress (data/collate.rs:197-349).

    edges: iterable of (row, col) in CONUS positions; gauges: list of CONUS
    gauge positions; parent_offset: the CONUS reach-subdivision map or None.
    Returns (active, rows, cols, gauge_compressed, outflow_idx), plus the
    compressed parent map when parent_offset is given.
    """
    edges = sorted(set(edges))
    active = sorted({r for r, _ in edges} | {c for _, c in edges} | set(gauges))
    if not active:
        raise ValueError("compress: empty active set")
    m = {p: k for k, p in enumerate(active)}
    rows = [m[r] for r, _ in edges]
    cols = [m[c] for _, c in edges]
    for r, c in zip(rows, cols):
        if r < c:
            raise ValueError("lower-triangular violated")
    gc = [m[g] for g in gauges]
    poff = None if parent_offset is None else compressed_parent_offset(active, list(parent_offset))
    if ddr_match:
        # the reference filters the edges by row for each gauge; grouping the
        # columns by row once, in edge order, gives the same lists
        by_row = {}
        for r, c in zip(rows, cols):
            by_row.setdefault(r, []).append(c)
        outflow = [list(by_row[g]) if g in by_row else [g] for g in gc]
    elif poff is not None:
        outflow = [[poff[_parent_of_row(poff, g) + 1] - 1] for g in gc]
    else:
        outflow = [[g] for g in gc]
    if parent_offset is None:
        return active, rows, cols, gc, outflow
    return active, rows, cols, gc, outflow, poff

"""Index construction at the workloads' sizes, bit-exact against the oracle
(oracle/topo.py, oracle.pattern_from_coo), run with the GPU tests so that the
batch the device routes and the checks of it travel together. Host work only.

* c5's batch: the 2,000 nested gauges' upstream subgraphs on the
  350,000-reach tree as a gauges store, their union (data/collate.rs:31-48)
  and compression with outflow_idx in both conventions (197-349), product
  against oracle; and the batch the workload builder routes equals it;
* the c2 network's CsrPattern (sparse/mod.rs:173-304) and the device plan's
  schedule (every upstream reach at an earlier position);
* topological order (adjacency/build.rs:170-272) of a 350,000-reach MERIT-style
  fabric in shuffled record order (at most four upstream COMIDs per reach, as
  the store holds), and the 1,000-graph rustworkx fixture.
"""
import json
import os
import sys

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import topo

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

from ddrs_amd import network as NW  # noqa: E402
from ddrs_amd import store as ST  # noqa: E402
from ddrs_amd import train as TR  # noqa: E402
from ddrs_amd import workloads as W  # noqa: E402
from ddrs_amd.sparse import Network  # noqa: E402


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def gauge_subgraphs(R, rows, cols, gauges):
    """Each gauge's upstream subgraph (the edges whose downstream reach lies in
    its basin), as the reference's gauges store holds them. A depth-first
    preorder numbers every basin as one interval [tin[g], tin[g] + size[g])."""
    size = TR.upstream_sizes(R, rows, cols)
    kids = [[] for _ in range(R)]
    for r, c in zip(rows.tolist(), cols.tolist()):
        kids[r].append(c)
    down = np.full(R, -1, np.int64)
    down[cols] = rows
    tin = np.zeros(R, np.int64)
    t = 0
    for root in np.flatnonzero(down < 0)[::-1]:
        stack = [int(root)]
        while stack:
            v = stack.pop()
            tin[v] = t
            t += 1
            stack.extend(kids[v])
    tr = tin[rows]
    out = {}
    for k, g in enumerate(gauges):
        keep = (tr >= tin[g]) & (tr < tin[g] + size[g])
        out[f"{k:08d}"] = (int(g), rows[keep].astype(np.int32), cols[keep].astype(np.int32))
    return out


def test_c5_batch_union_and_compress():
    cfg = W.CONFIGS["c5"]
    R0 = cfg["R"]
    rows, cols = NW.gen_tree(R0, cfg["window"], 1005)
    g_full = TR.nested_gauges(R0, rows, cols, cfg["gauges"], cfg["size_lo"], cfg["size_hi"],
                              seed=7005)
    say(f"c5 index: {g_full.size} gauges on {R0} reaches")
    subs = gauge_subgraphs(R0, rows, cols, g_full)
    staids = list(subs) + ["99999999"]          # a STAID the store lacks is skipped
    store = ST.GagesAdjacencyStore("<c5>", {s: ST.GageSubgraph(s, g, "", r, c)
                                           for s, (g, r, c) in subs.items()})
    u = ST.union_subgraphs(staids, store)
    we, wg = topo.union_subgraphs(staids, {s: (g, "", list(zip(r.tolist(), c.tolist())))
                                           for s, (g, r, c) in subs.items()})
    assert list(zip(u.rows.tolist(), u.cols.tolist())) == we
    assert [(s, g) for s, g, _ in u.gauges] == [(s, g) for s, g, _ in wg]
    say(f"union: {len(we)} edges")
    # the workload builder's union of the same gauges is this union
    ur, uc = TR.upstream_union(R0, rows, cols, g_full)
    order = np.lexsort((uc, ur))
    assert np.array_equal(ur[order], u.rows) and np.array_equal(uc[order], u.cols)
    for ddr_match in (False, True):
        c = NW.compress_ex(u.rows, u.cols, u.gauge_rows(), ddr_match)
        a, r, cc, gc, of = topo.compress(we, [g for _, g, _ in wg], ddr_match)
        assert c.active.tolist() == a and c.rows.tolist() == r and c.cols.tolist() == cc
        assert c.gauge_compressed.tolist() == gc
        assert [x.tolist() for x in c.outflow_idx()] == of
    wl = W.build_c5()
    assert np.array_equal(wl.rows, c.rows) and np.array_equal(wl.cols, c.cols)
    assert np.array_equal(wl.gauges, c.gauge_compressed)


def test_c2_csr_pattern_and_schedule():
    wl = W.build("c2")
    say(f"c2 CsrPattern: {wl.R} reaches")
    net = Network(wl.adjacency())
    got = net.csr()
    want = O.pattern_from_coo(wl.R, wl.rows, wl.cols)
    for f in ("crow", "col", "row_for_nnz", "adj_values", "diag_mask", "trans_crow",
              "trans_col", "trans_to_orig"):
        assert np.array_equal(got[f], getattr(want, f)), f
    perm, level = net.schedule()
    pos = np.empty(wl.R, np.int64)
    pos[perm[perm >= 0]] = np.flatnonzero(perm >= 0)
    assert np.all(pos[wl.cols] < pos[wl.rows])      # upstream first
    assert np.all(level[wl.rows] > level[wl.cols])


def merit_records(R, seed):
    """A MERIT-style fabric: reach i drains to one of the next 256 with room
    for another upstream (at most four per reach); COMIDs a seeded
    permutation; records in shuffled order."""
    rng = np.random.default_rng(seed)
    ups = [[] for _ in range(R)]
    for i in range(R - 1):
        hi = min(i + 256, R - 1)
        for _ in range(8):
            j = int(rng.integers(i + 1, hi + 1))
            if len(ups[j]) < 4:
                ups[j].append(i)
                break
    comid = rng.permutation(R).astype(np.int64) + 70_000_000
    rec = np.zeros((R, 5), np.int64)
    rec[:, 0] = comid
    for j in range(R):
        for k, u in enumerate(ups[j]):
            rec[j, 1 + k] = comid[u]
    return rec[rng.permutation(R)]


def test_topology_at_c2_size_and_fuzz(golden_dir):
    R = 350_000
    say(f"topology: {R} reaches")
    rec = merit_records(R, 4101)
    order, rows, cols, dropped = NW.build_topology(rec)
    worder, wrows, wcols, wdropped = topo.build_topo_and_coo(
        [(int(r[0]), [int(u) for u in r[1:] if u > 0]) for r in rec])
    assert order.tolist() == worder and dropped.tolist() == wdropped
    assert rows.tolist() == wrows and cols.tolist() == wcols
    n = 0
    with open(os.path.join(golden_dir, "toposort_fuzz.jsonl")) as fh:
        for line in fh:
            if line.strip():
                case = json.loads(line)
                r = np.array([x + [0] * (5 - len(x)) for x in case["records"]], np.int64)
                assert NW.build_topology(r)[0].tolist() == case["order"]
                n += 1
    assert n == 1000

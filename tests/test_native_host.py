"""Native host-side work of the HIP library, checked on the CPU: the library
loads and exports every symbol of include/ddrs_hip.h; index construction is
bit-exact against the oracle and the reference's fixtures. No GPU calls."""
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from oracle import topo
from tests.nets import chain, random_tree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    from ddrs_amd import _native
    header = open(os.path.join(ROOT, "include", "ddrs_hip.h")).read()
    declared = set(re.findall(r"\b(ddrs_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 25
    for name in sorted(declared):
        assert hasattr(_native.lib, name), name
    assert declared <= set(_native.EXPORTED)


def _plan_csr(R, rows, cols, vals=None):
    from ddrs_amd.sparse import Network, SparseAdjacency
    net = Network(SparseAdjacency.from_coo(R, rows, cols, vals))
    return net, net.csr()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_csr_pattern_bit_exact(seed):
    """CsrPattern::from_sparse + transposed view (sparse/mod.rs:173-304), product
    builder against the oracle's, including unsorted input and non-unit weights."""
    rng = np.random.default_rng(seed)
    R = 2000
    rows, cols = random_tree(R, 64, rng)
    perm = rng.permutation(rows.shape[0])
    rows, cols = rows[perm], cols[perm]
    vals = rng.uniform(0.5, 1.5, rows.shape[0]).astype(np.float32)
    _, got = _plan_csr(R, rows, cols, vals)
    want = O.pattern_from_coo(R, rows, cols, vals)
    for k in ["crow", "col", "row_for_nnz", "adj_values", "diag_mask", "trans_crow",
              "trans_col", "trans_to_orig"]:
        np.testing.assert_array_equal(got[k], getattr(want, k), err_msg=k)


def test_schedule_is_level_major_topological():
    rng = np.random.default_rng(5)
    R = 5000
    rows, cols = random_tree(R, 256, rng)
    net, _ = _plan_csr(R, rows, cols)
    perm, level = net.schedule()
    assert sorted(perm.tolist()) == list(range(R))
    pos = np.empty(R, np.int64)
    pos[perm] = np.arange(R)
    # every upstream reach sits at an earlier schedule position, one level lower at least
    assert (pos[cols] < pos[rows]).all()
    assert (level[cols] < level[rows]).all()
    # longest upstream path, computed independently
    lev = np.zeros(R, np.int64)
    order = np.argsort(cols, kind="stable")
    for j in order:
        lev[rows[j]] = max(lev[rows[j]], lev[cols[j]] + 1)
    np.testing.assert_array_equal(level, lev)
    assert np.all(np.diff(level[perm]) >= 0)
    assert net.n_levels == lev.max() + 1


def test_forest_of_small_basins_fits_chunk_staging():
    """Small basins put outlets with several direct tributaries on deeper
    levels; the schedule closes a chunk before its upstream links overflow the
    192-link staging (a plan that used to be refused with DDRS_ERR_CAPACITY)."""
    from ddrs_amd import network as NW
    rng = np.random.default_rng(9)
    sizes = np.minimum(rng.zipf(1.8, 2000), 3000).astype(np.int64)
    rows, cols, outlets = NW.gen_forest(sizes, 256, 77)
    # in plain level-major order some 64 consecutive reaches hold > 192 upstream links
    indeg = np.bincount(rows, minlength=int(sizes.sum()))
    lev = np.zeros(int(sizes.sum()), np.int64)
    for j in np.argsort(rows, kind="stable"):
        lev[rows[j]] = max(lev[rows[j]], lev[cols[j]] + 1)
    lm = np.argsort(lev, kind="stable")
    assert max(indeg[lm[k:k + 64]].sum() for k in range(0, lm.size, 64)) > 192
    R = int(sizes.sum())
    net, _ = _plan_csr(R, rows, cols)
    perm, level = net.schedule()
    pos = np.empty(R, np.int64)
    pos[perm] = np.arange(R)
    assert (pos[cols] < pos[rows]).all() and (level[cols] < level[rows]).all()
    assert net.n_chunks >= int(np.ceil(rows.size / 192))
    assert (outlets == np.cumsum(sizes) - 1).all()


def test_plan_rejects_bad_adjacency():
    from ddrs_amd.sparse import Network, SparseAdjacency
    from ddrs_amd._native import DdrsError
    with pytest.raises(DdrsError):
        Network(SparseAdjacency.from_coo(3, [0], [1]))      # upper triangular
    with pytest.raises(DdrsError):
        Network(SparseAdjacency.from_coo(3, [5], [1]))      # out of range
    with pytest.raises(DdrsError):
        Network(SparseAdjacency.from_coo(0, [], []))        # empty network


def test_topology_matches_rustworkx(golden_dir):
    """Native build_topology against the 1000-graph rustworkx fixture
    (tests/adjacency_build.rs:191-232)."""
    from ddrs_amd.network import build_topology
    n = 0
    with open(os.path.join(golden_dir, "toposort_fuzz.jsonl")) as fh:
        for line in fh:
            if not line.strip():
                continue
            case = json.loads(line)
            rec = np.array([r + [0] * (5 - len(r)) for r in case["records"]], np.int64)
            order, rows, cols, _ = build_topology(rec)
            assert order.tolist() == case["order"]
            _, orows, ocols, _ = topo.build_topo_and_coo(
                [(r[0], [u for u in r[1:] if u > 0]) for r in case["records"]])
            assert rows.tolist() == orows and cols.tolist() == ocols
            n += 1
    assert n >= 1000


def test_topology_cycles_and_non_dendritic():
    from ddrs_amd.network import build_topology
    from ddrs_amd._native import DdrsError
    rec = np.array([[1, 2, 0, 0, 0], [2, 1, 0, 0, 0], [3, 1, 0, 0, 0], [4, 3, 0, 0, 0],
                    [5, 0, 0, 0, 0], [6, 5, 0, 0, 0]], np.int64)
    order, rows, cols, dropped = build_topology(rec)
    want = topo.build_topo_and_coo([(r[0], [u for u in r[1:] if u > 0]) for r in rec.tolist()])
    assert order.tolist() == want[0] and dropped.tolist() == want[3] == [1, 2]
    with pytest.raises(DdrsError):
        build_topology(np.array([[10, 1, 0, 0, 0], [11, 1, 0, 0, 0]], np.int64))


@pytest.mark.parametrize("seed", [0, 1])
def test_compress_matches_oracle(seed):
    """collate::compress (data/collate.rs:197-349) on a union of gauge subgraphs."""
    from ddrs_amd.network import compress
    rng = np.random.default_rng(seed)
    R = 3000
    rows, cols = random_tree(R, 32, rng)
    down = np.full(R, -1)
    down[cols] = rows
    gauges = rng.choice(R, 20, replace=False)
    er, ec = [], []
    for g in gauges:  # every edge of the basin above the gauge
        basin = {int(g)}
        for j in range(int(g) - 1, -1, -1):
            if down[j] in basin:
                basin.add(j)
        for j in basin:
            if j != g:
                er.append(int(down[j]))
                ec.append(j)
    active, r2, c2, gc = compress(er, ec, gauges)
    wa, wr, wc, wg, _ = topo.compress(list(zip(er, ec)), gauges.tolist())
    assert active.tolist() == wa and r2.tolist() == wr and c2.tolist() == wc and gc.tolist() == wg


def test_generators_are_deterministic():
    from ddrs_amd import network as NW
    a = NW.gen_forcing(50, 30, seed=9, storms=True)
    b = NW.gen_forcing(50, 30, seed=9, storms=True)
    np.testing.assert_array_equal(a, b)
    assert (a > 0).all()
    r1, c1 = NW.gen_tree(1000, 256, 3)
    assert (r1 > c1).all() and (r1 - c1 <= 256).all()
    at = NW.gen_attributes(1000, 4)
    assert at["length"].min() >= 1000 and at["length"].max() <= 10000
    assert at["slope"].min() >= 1e-4 and at["slope"].max() <= 5e-2


def test_workspace_size_query_and_window_choice():
    """ddrs_workspace_query_bytes (no allocation) and the autotuner's
    checkpoint-window choice: the longest window that fits the budget, all
    steps when they fit, an error when one step does not."""
    from ddrs_amd import network as NW
    from ddrs_amd import tune as TU
    from ddrs_amd.sparse import Network, SparseAdjacency
    R = 20000
    rows, cols = NW.gen_tree(R, 256, 3)
    net = Network(SparseAdjacency.from_coo(R, rows, cols))
    b = lambda w, m, d=True: net.workspace_bytes(w, m, d)
    # the dense dL/dQ buffer is max_steps x positions x 4 bytes
    P = (b(10, 100) - b(10, 100, False)) // (100 * 4)
    assert R <= P < R + 64 * net.n_chunks
    # per window step: 68 bytes per position (history, sums, messages, records, gb)
    assert b(11, 100) - b(10, 100) == 68 * P
    T = 500
    full = b(T - 1, T - 1, False)
    assert TU.choose_window(net, T, full) == T - 1
    w = TU.choose_window(net, T, full - 1)
    assert w < T - 1 and b(w, T - 1, False) <= full - 1 < b(w + 1, T - 1, False)
    # the size is not monotone in the window: checkpoints grow as it shrinks
    assert b(1, T - 1, False) > b(2, T - 1, False)
    assert TU.choose_window(net, T, b(2, T - 1, False)) >= 2
    with pytest.raises(MemoryError):
        TU.choose_window(net, T, min(b(w, T - 1, False) for w in range(1, T)) - 1)

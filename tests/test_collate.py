"""Batch collation: union_subgraphs and compress with outflow_idx and the
reach-subdivision map (data/collate.rs:31-48, 141-180, 197-349), product
(ddrs_amd.store.union_subgraphs, ddrs_amd.network.compress_ex -> the C ABI's
ddrs_compress_ex) and oracle (oracle/topo.py) alike.

Pins: the reference's own known answers (collate.rs:443-719, restated below
with their expected values) for both implementations, then the two compared
bit for bit on random subdivided networks, both outflow conventions."""
import os

import numpy as np
import pytest

from oracle import topo
from ddrs_amd import _native as NAT
from ddrs_amd import network as NW
from ddrs_amd import store as ST


def product_union(gauges, staids):
    subs = {gid: ST.GageSubgraph(gid, gi, f"comid{gi}",
                                 np.array([r for r, _ in e], np.int32),
                                 np.array([c for _, c in e], np.int32))
            for gid, gi, e in gauges}
    u = ST.union_subgraphs(staids, ST.GagesAdjacencyStore("<inline>", subs))
    return list(zip(u.rows.tolist(), u.cols.tolist())), u.gauges


def oracle_union(gauges, staids):
    return topo.union_subgraphs(staids, {gid: (gi, f"comid{gi}", e) for gid, gi, e in gauges})


@pytest.mark.parametrize("union", [product_union, oracle_union])
def test_union_known_answers(union):
    # collate.rs:443-462: shared edge (2, 1) once, sorted by (row, col)
    g = [("0000000A", 3, [(3, 1), (3, 2), (2, 1)]), ("0000000B", 4, [(4, 2), (2, 1)])]
    edges, gauges = union(g, ["0000000A", "0000000B"])
    assert edges == [(2, 1), (3, 1), (3, 2), (4, 2)]
    assert gauges == [("0000000A", 3, "comid3"), ("0000000B", 4, "comid4")]
    # 464-472: a missing STAID is skipped
    edges, gauges = union([("0000000A", 3, [(3, 1)])], ["0000000A", "00000099"])
    assert edges == [(3, 1)] and [x[0] for x in gauges] == ["0000000A"]
    # 474-482: an empty batch
    assert union([("0000000A", 3, [(3, 1)])], []) == ([], [])


def product_compress(edges, gauges, ddr_match, parent_offset=None):
    e = np.asarray(edges, np.int32).reshape(-1, 2)
    c = NW.compress_ex(e[:, 0], e[:, 1], gauges, ddr_match, parent_offset)
    out = (c.active.tolist(), c.rows.tolist(), c.cols.tolist(), c.gauge_compressed.tolist(),
           [x.tolist() for x in c.outflow_idx()])
    return out if parent_offset is None else out + (c.parent_offset.tolist(),)


def oracle_compress(edges, gauges, ddr_match, parent_offset=None):
    return topo.compress(edges, gauges, ddr_match, parent_offset)


IMPLS = [product_compress, oracle_compress]


@pytest.mark.parametrize("compress", IMPLS)
def test_compress_known_answers(compress):
    # collate.rs:485-510: topological order kept; ddr_match = the upstream cols
    a, r, c, g, out = compress([(2, 0), (3, 1), (4, 2), (4, 3)], [4, 3], True)
    assert (a, r, c, g, out) == ([0, 1, 2, 3, 4], [2, 3, 4, 4], [0, 1, 2, 3], [4, 3],
                                 [[2, 3], [1]])
    # 512-529: a sparse active set remapped densely, lower-triangular
    a, r, c, g, _ = compress([(9, 7), (9, 5), (7, 2)], [9], True)
    assert a == [2, 5, 7, 9] and g == [3] and all(x >= y for x, y in zip(r, c))
    # 561-599: without ddr_match the gauge's own reach, with it the upstream cols
    e = [(2, 0), (2, 1)]
    assert compress(e, [2], False)[4] == [[2]]
    assert compress(e, [2], True)[4] == [[0, 1]]
    # 601-619: no incoming edges: both conventions give the gauge itself
    assert compress([], [2], True)[3:] == ([0], [[0]])
    assert compress([], [2], False)[3:] == ([0], [[0]])
    # 627-671: the subdivided gauge reach (parents 0, 1 single pieces; parent 2
    # split into rows 2..5): the map renumbered, outflow at the outlet piece
    sub = [(2, 0), (2, 1), (3, 2), (4, 3), (5, 4)]
    res = compress(sub, [5], False, [0, 1, 2, 6])
    assert res[5] == [0, 1, 2, 6] and res[3] == [5] and res[4] == [[5]]
    # 696-719: the identity map leaves outflow at the gauge reach
    res = compress(e, [2], False, [0, 1, 2, 3])
    assert res[5] == [0, 1, 2, 3] and res[4] == [[2]] and res[4] == compress(e, [2], False)[4]


@pytest.mark.parametrize("compress", IMPLS)
def test_compress_rejects_bad_input(compress):
    err = (NAT.DdrsError, ValueError)
    # 531-545: an edge that is not lower triangular
    with pytest.raises(err):
        compress([(0, 1)], [0], True)
    # 547-559: an empty active set
    with pytest.raises(err):
        compress([], [], True)
    # 673-694: the gauge reach's inlet piece missing (a partial sub-reach chain)
    with pytest.raises(err, match="contiguous"):
        compress([(4, 3), (5, 4)], [5], False, [0, 1, 2, 6])


def subdivided(n_parent, seed):
    """A tree of parents, each split into 1-4 pieces chained in order; the
    outlet piece of a parent drains to the first piece of its downstream
    parent. Returns (parent_offset, rows, cols, down parent)."""
    rng = np.random.default_rng(seed)
    prow, pcol = NW.gen_tree(n_parent, 16, seed)
    down = np.full(n_parent, -1)
    down[pcol] = prow
    m = rng.integers(1, 5, n_parent)
    off = np.concatenate([[0], np.cumsum(m)]).astype(np.int32)
    rows, cols = [], []
    for p in range(n_parent):
        for k in range(m[p] - 1):
            rows.append(off[p] + k + 1)
            cols.append(off[p] + k)
        if down[p] >= 0:
            rows.append(off[down[p]])
            cols.append(off[p + 1] - 1)
    return off, np.array(rows, np.int32), np.array(cols, np.int32), down


def gauge_subgraph(off, rows, cols, down, g):
    """Every edge whose downstream row lies in a parent upstream of (or at)
    parent g; gage_idx the outlet piece of g."""
    up = {g}
    changed = True
    while changed:
        changed = False
        for p in np.flatnonzero(np.isin(down, list(up))):
            if p not in up:
                up.add(int(p))
                changed = True
    keep_rows = np.zeros(off[-1], bool)
    for p in up:
        keep_rows[off[p]:off[p + 1]] = True
    k = keep_rows[rows]
    return int(off[g + 1] - 1), list(zip(rows[k].tolist(), cols[k].tolist()))


@pytest.mark.parametrize("seed", range(6))
def test_product_equals_oracle_on_subdivided_batches(seed):
    off, rows, cols, down = subdivided(300, 40 + seed)
    rng = np.random.default_rng(seed)
    gauges = []
    for k, g in enumerate(rng.choice(300, 12, replace=False)):
        gi, e = gauge_subgraph(off, rows, cols, down, int(g))
        gauges.append((f"G{k:07d}", gi, e))
    staids = [g[0] for g in gauges] + ["MISSING0"]
    eu, gu = product_union(gauges, staids)
    assert (eu, gu) == oracle_union(gauges, staids)
    grow = [g[1] for g in gu]
    for ddr_match in (False, True):
        for po in (None, off):
            got = product_compress(eu, grow, ddr_match, po)
            want = oracle_compress(eu, grow, ddr_match, None if po is None else po.tolist())
            assert got == tuple(want) if isinstance(want, tuple) else got == want


def test_union_over_the_juniata_gauges_store(golden_dir):
    """The reference's Juniata stores: the product union (and its compress)
    equal the oracle's, with a STAID the store lacks in the request."""
    base = os.path.join(golden_dir, "juniata")
    conus = ST.ConusAdjacencyStore.open(os.path.join(base, "juniata_conus_adjacency.zarr"))
    staids = ["01567000", "99999999"]
    gages = ST.GagesAdjacencyStore.open(os.path.join(base, "juniata_gages_adjacency.zarr"), staids)
    u = ST.union_subgraphs(staids, gages)
    g = gages.get("01567000")
    want_e, want_g = topo.union_subgraphs(
        staids, {"01567000": (g.gage_idx, g.gage_catchment,
                              list(zip(g.indices_0.tolist(), g.indices_1.tolist())))})
    assert list(zip(u.rows.tolist(), u.cols.tolist())) == want_e and u.gauges == want_g
    for ddr_match in (False, True):
        c = NW.compress_ex(u.rows, u.cols, u.gauge_rows(), ddr_match, conus.parent_offset)
        w = topo.compress(want_e, [x[1] for x in want_g], ddr_match, conus.parent_offset.tolist())
        assert (c.active.tolist(), c.rows.tolist(), c.cols.tolist()) == tuple(w[:3])
        assert [x.tolist() for x in c.outflow_idx()] == w[4] and c.parent_offset.tolist() == w[5]

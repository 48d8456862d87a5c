"""Pins the CPU oracle against the golden vectors the reference's own tests
hold. CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import topo
from tests.nets import (GOLDEN_X_SOL_BITS, POS_LENGTH, POS_QT, chain, stress_params)


def test_golden_x_sol_bits_bit_exact():
    """tests/positivity_clamp.rs:59-70,372-396: the raw S27 solve of the 10-reach
    stress fixture with enforce_positivity off, bit for bit (f32)."""
    R = 10
    pat = O.pattern_from_coo(R, *chain(R))
    P = stress_params(O, False)
    _, saved, _ = O.step_fwd(P, pat, np.full(R, 0.035), np.full(R, 0.4), np.full(R, 20.0),
                             POS_QT, np.full(R, 0.01), POS_LENGTH, np.full(R, 0.001))
    bits = saved[O.SAVE["x_sol"]].astype(np.float32).view(np.uint32)
    assert [int(b) for b in bits] == GOLDEN_X_SOL_BITS


def test_golden_positivity_clamp_moves_the_solve():
    """Non-vacuity of the golden test (positivity_clamp.rs:389-395): with the clamp
    on, the fixture's solve changes and has no negatives (317-331)."""
    R = 10
    pat = O.pattern_from_coo(R, *chain(R))
    args = (np.full(R, 0.035), np.full(R, 0.4), np.full(R, 20.0), POS_QT, np.full(R, 0.01),
            POS_LENGTH, np.full(R, 0.001))
    _, off, _ = O.step_fwd(stress_params(O, False), pat, *args)
    _, on, _ = O.step_fwd(stress_params(O, True), pat, *args)
    x_off, x_on = off[O.SAVE["x_sol"]], on[O.SAVE["x_sol"]]
    assert (x_off < 0).sum() > 0
    assert (x_on < 0).sum() == 0
    assert np.any(x_on != x_off)
    assert on[O.SAVE["c1"]].min() >= 0 and on[O.SAVE["c3"]].min() >= 0
    assert off[O.SAVE["c1"]].min() < 0 and off[O.SAVE["c3"]].min() < 0


def test_hotstart_fixture_77006074(golden_dir):
    """tests/training_step_layer_b.rs:242-350: DDR's cold start for gauge
    10336740 (19 reaches), (I - N) Q0 = q'_0 then clamp, within 1e-5."""
    sg = np.load(os.path.join(golden_dir, "subgraph_77006074.npz"), allow_pickle=False)
    hs = np.load(os.path.join(golden_dir, "hotstart_77006074.npz"), allow_pickle=False)
    n = hs["hotstart"].shape[0]
    pat = O.pattern_from_coo(n, sg["rows"], sg["cols"], sg["vals"])
    q0 = O.hotstart(O.make_params(), pat, hs["q_prime_t0"])
    np.testing.assert_allclose(q0, hs["hotstart"], rtol=1e-5, atol=1e-5)
    # non-vacuity: the network really accumulates (an identity solve fails)
    assert np.abs(hs["hotstart"] - np.maximum(hs["q_prime_t0"], 1e-4)).max() > 1e-3


def test_toposort_matches_rustworkx(golden_dir):
    """tests/adjacency_build.rs:191-232: 1000 random dendritic networks, order
    element for element against rustworkx 0.17.1."""
    total = 0
    with open(os.path.join(golden_dir, "toposort_fuzz.jsonl")) as fh:
        for line in fh:
            if not line.strip():
                continue
            case = json.loads(line)
            recs = [(r[0], [u for u in r[1:] if u > 0]) for r in case["records"]]
            order, rows, cols, _ = topo.build_topo_and_coo(recs)
            assert order == case["order"]
            assert all(r >= c for r, c in zip(rows, cols))
            total += 1
    assert total >= 1000


def test_toposort_drops_cycles():
    """build.rs:238-262: records of COMIDs on a simple cycle are dropped and the
    build retried; an upstream reference to a dropped COMID still makes it a node
    (only records are filtered)."""
    recs = [(1, [2]), (2, [1]), (3, [1]), (4, [3]), (5, []), (6, [5])]
    order, rows, cols, dropped = topo.build_topo_and_coo(recs)
    assert dropped == [1, 2]
    assert set(order) == {1, 3, 4, 5, 6}
    assert 2 not in order


def test_pattern_layout():
    """CsrPattern::from_sparse (sparse/mod.rs:173-227): off-diagonals ascending
    by column then the diagonal; transposed view (276-304)."""
    rows = np.array([3, 2, 3, 4], np.int32)
    cols = np.array([2, 0, 1, 3], np.int32)
    pat = O.pattern_from_coo(5, rows, cols)
    assert pat.crow.tolist() == [0, 1, 2, 4, 7, 9]
    assert pat.col.tolist() == [0, 1, 0, 2, 1, 2, 3, 3, 4]
    assert pat.diag_mask.tolist() == [1, 1, 0, 1, 0, 0, 1, 0, 1]
    assert pat.trans_crow.tolist() == [0, 2, 4, 6, 8, 9]
    assert pat.trans_col.tolist() == [0, 2, 1, 3, 2, 3, 3, 4, 4]
    with pytest.raises(ValueError):
        O.pattern_from_coo(3, np.array([0], np.int32), np.array([1], np.int32))


def test_compress_invariants():
    """collate::compress (data/collate.rs:197-349): the active set is sorted and
    unique; every compressed edge and gauge maps back to its CONUS position;
    lower-triangular is kept; ddr_match reads the gauge's upstream columns."""
    rng = np.random.default_rng(3)
    R = 400
    down = np.array([i + 1 + rng.integers(0, min(8, R - 1 - i)) for i in range(R - 1)])
    edges = [(int(down[i]), i) for i in range(R - 1) if rng.random() < 0.6]
    gauges = sorted(int(g) for g in rng.choice(R, 7, replace=False))
    active, rows, cols, gc, out = topo.compress(edges + edges[:5], gauges)
    assert active == sorted(set(active))
    assert sorted(set(edges)) == [(active[r], active[c]) for r, c in zip(rows, cols)]
    assert [active[g] for g in gc] == gauges
    assert all(r >= c for r, c in zip(rows, cols))
    assert out == [[g] for g in gc]
    _, rows2, cols2, gc2, out2 = topo.compress(edges, gauges, ddr_match=True)
    for g, o in zip(gc2, out2):
        ups = [c for r, c in zip(rows2, cols2) if r == g]
        assert o == (ups if ups else [g])

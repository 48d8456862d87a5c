"""The autotuner's arithmetic on the CPU (ddrs_amd.tune.tune_schedule), with
its GPU probe replaced by a timing model of known lines: a fill term plus a
per-step term for forward and forward + adjoint, per sub-level cap. Pins the
direct mode (a short window timed at its own length), the extrapolation
through two probe lengths, the recomputed forward segments added when the
checkpoint window is shorter than the window, and the choice of the smallest
prediction. The plans are built on the host; nothing runs on a GPU."""
import numpy as np
import pytest

from tests.nets import random_tree

# cap -> (fwd fill ms, fwd ms/step, total fill ms, total ms/step)
MODEL = {2: (5.0, 0.010, 9.0, 0.030), 4: (2.0, 0.012, 4.0, 0.034), 8: (1.0, 0.014, 2.0, 0.036)}


@pytest.fixture
def tuner(monkeypatch):
    from ddrs_amd import tune as TU
    calls = []

    def fake_probe(net, params, adj, lengths, window, *args):
        f0, fs, t0, ts = MODEL[net.max_sub_levels]
        calls.append((net.max_sub_levels, tuple(lengths), window))
        return [(f0 + fs * L, t0 + ts * L) for L in lengths]

    monkeypatch.setattr(TU, "_probe", fake_probe)
    return TU, calls


def _adj():
    from ddrs_amd.sparse import SparseAdjacency
    R = 3000
    rows, cols = random_tree(R, 64, np.random.default_rng(5))
    return SparseAdjacency.from_coo(R, rows, cols)


def test_direct_mode_times_the_window_itself(tuner):
    TU, calls = tuner
    T = 301                                  # 300 steps <= 2.5 x 288
    ts = TU.tune_schedule(_adj(), T, budget_bytes=1 << 40)
    assert [c[1] for c in calls] == [(300,)] * 3
    assert ts.window == 300
    for p in ts.probes:
        f0, fs, t0, tsm = MODEL[p["sub_levels"]]
        assert p["predicted_ms"] == pytest.approx(t0 + tsm * 300)
    # 2: 18.0, 4: 14.2, 8: 12.8
    assert ts.max_sub_levels == 8 and ts.predicted_ms == pytest.approx(12.8)
    assert ts.network.max_sub_levels == 8


def test_extrapolation_and_choice(tuner):
    TU, calls = tuner
    T = 8760
    ts = TU.tune_schedule(_adj(), T, budget_bytes=1 << 40, probe_steps=(96, 288))
    assert all(c[1] == (96, 288) for c in calls)
    steps = T - 1
    want = {sl: t0 + tsm * steps for sl, (f0, fs, t0, tsm) in MODEL.items()}
    for p in ts.probes:
        assert p["predicted_ms"] == pytest.approx(want[p["sub_levels"]], rel=1e-9)
        assert p["total_ms_per_step"] == pytest.approx(MODEL[p["sub_levels"]][3])
        assert p["fwd_ms_per_step"] == pytest.approx(MODEL[p["sub_levels"]][1])
    # long windows favour the smallest per-step term: cap 2
    assert ts.max_sub_levels == 2
    assert ts.predicted_ms == pytest.approx(min(want.values()))


def test_recomputed_segments_are_added(tuner):
    TU, calls = tuner
    from ddrs_amd.sparse import Network
    adj = _adj()
    T = 2001
    steps = T - 1
    # a budget that fits a 700-step checkpoint window at cap 2 and no longer one
    budget = Network(adj, 2, 64).workspace_bytes(700, steps, False)
    ts = TU.tune_schedule(adj, T, budget_bytes=budget, sub_levels=(2,), probe_steps=(96, 288))
    w = ts.window
    assert w == TU.choose_window(Network(adj, 2, 64), T, budget) and w < steps
    f0, fs, t0, tsm = MODEL[2]
    last = steps - w * ((steps - 1) // w)
    assert 1 <= last <= w
    want = t0 + tsm * steps + fs * (steps - last)
    assert ts.predicted_ms == pytest.approx(want, rel=1e-9)
    assert calls[0][2] == w

"""The schedule autotuner on the GPU (ddrs_amd.tune): it returns a layout
from its candidates with a finite prediction, in both the direct mode (the
workload's own length) and the extrapolating mode (two probe lengths), and
routing with the tuned network and window gives the default layout's bits
(results do not depend on the schedule or the checkpoint window). Needs an
MI355X."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

import ddrs_amd as D  # noqa: E402
from ddrs_amd import network as NW  # noqa: E402
from ddrs_amd import tune as TU  # noqa: E402
from ddrs_amd.sparse import Network, SparseAdjacency  # noqa: E402

DEV = "cuda:0"


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def t(x):
    return torch.as_tensor(np.ascontiguousarray(x, np.float32), device=DEV)


def _route(net, at, qp, q0, gauges, window):
    n = t(at["n"]).requires_grad_(True)
    qs = t(at["q_spatial"]).requires_grad_(True)
    p = torch.full((net.n,), 21.0, device=DEV)
    slope = torch.clamp_min(t(at["slope"]), 1e-3)
    Qg = D.route_window_gauged(net, D.Params(), n, qs, p, qp, q0, t(at["length"]), slope,
                               gauges, window=window)
    Qg.backward(torch.ones_like(Qg))
    torch.cuda.synchronize()
    return Qg.detach().cpu().numpy(), n.grad.cpu().numpy(), qs.grad.cpu().numpy()


@pytest.mark.parametrize("T,probe", [(400, (96, 288)), (1500, (48, 144))])
def test_tuned_schedule_routes_identically(T, probe):
    R = 30000
    say(f"autotuner R={R} T={T} probes {probe}")
    rows, cols = NW.gen_tree(R, 256, 71)
    at = NW.gen_attributes(R, 72)
    adj = SparseAdjacency.from_coo(R, rows, cols, None, at["length"], at["slope"])
    qp = NW.gen_forcing_device(R, T, 73, True, device=DEV)
    gauges = np.array([R // 3, R // 2, R - 1], np.int32)
    ts = TU.tune_schedule(adj, T, q_prime=qp, gauges=gauges, sub_levels=(2, 8),
                          probe_steps=probe)
    say("chosen", ts.max_sub_levels, ts.block, ts.window, f"{ts.predicted_ms:.2f} ms",
        [(p["sub_levels"], round(p["predicted_ms"], 2)) for p in ts.probes])
    assert ts.max_sub_levels in (2, 8) and len(ts.probes) == 2
    assert ts.window == T - 1
    assert all(np.isfinite(p["predicted_ms"]) and p["predicted_ms"] > 0 for p in ts.probes)
    assert ts.probes[0]["probe_lengths"] == ([T - 1] if T - 1 <= 2.5 * max(probe)
                                             else list(probe))
    base = Network(adj)
    q0 = D.hotstart(base, D.Params(), qp[0])
    want = _route(base, at, qp, q0, gauges, None)
    got = _route(ts.network, at, qp, q0, gauges, None)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    # a budget that forces checkpoint segments: the window shrinks, the bits stay
    # (a fresh plan of the same layout: the tuned one already holds a
    # whole-window workspace, which a shorter window would reuse)
    budget = ts.network.workspace_bytes(T // 3, T - 1, False)
    w = TU.choose_window(ts.network, T, budget)
    assert w < T - 1
    seg = _route(Network(adj, ts.max_sub_levels, ts.block), at, qp, q0, gauges, w)
    for a, b in zip(seg, want):
        np.testing.assert_array_equal(a, b)

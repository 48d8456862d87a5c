"""Capture of the gauged window pair (persistent forward + adjoint over
checkpoint segments) in a HIP graph through torch.cuda.graph, replayed for
several forcings. Each persistent launch begins with a one-thread kernel that
sets its claim base and epoch on the device, so that a replay uses fresh
ones. Every replay, with eager work and synchronisations between replays,
must give the bits of the eager calls on the same inputs. Needs an MI355X."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

import ddrs_amd as D  # noqa: E402
from ddrs_amd import _native as NAT  # noqa: E402
from ddrs_amd import network as NW  # noqa: E402
from ddrs_amd.sparse import Network, SparseAdjacency, Workspace  # noqa: E402

DEV = "cuda:0"


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_captured_window_pair_replays_bitwise():
    R, T, window = 20000, 97, 32
    say(f"graph capture of the gauged window pair, R={R} T={T} window={window}")
    rows, cols = NW.gen_tree(R, 256, 91)
    at = NW.gen_attributes(R, 92)
    net = Network(SparseAdjacency.from_coo(R, rows, cols, None, at["length"], at["slope"]))
    ws = Workspace(net, window, T - 1, dense_grad=False)
    gauges = np.sort(np.concatenate([[R - 1], np.random.default_rng(93).choice(R - 1, 15,
                                                                                replace=False)]))
    ws.set_gauges(gauges.astype(np.int32))
    K = gauges.size
    f = lambda x: torch.as_tensor(np.ascontiguousarray(x, np.float32), device=DEV)
    params = D.Params()
    cp = params.to_c()
    s0 = torch.cuda.current_stream().cuda_stream
    NAT.check(NAT.lib.ddrs_mc_bind_statics(ws.handle, cp, f(at["n"]).data_ptr(),
                                           f(at["q_spatial"]).data_ptr(),
                                           torch.full((R,), 21.0, device=DEV).data_ptr(),
                                           f(at["length"]).data_ptr(),
                                           torch.clamp_min(f(at["slope"]), 1e-3).data_ptr(),
                                           None, s0))
    # static buffers the graph reads and writes
    qp = torch.empty((T, R), device=DEV)
    q0 = torch.empty(R, device=DEV)
    gQg = torch.empty((T, K), device=DEV)
    Qg = torch.empty((T, K), device=DEV)
    gn, gq = torch.empty(R, device=DEV), torch.empty(R, device=DEV)
    ptr = lambda x: x.data_ptr()

    def pair(stream):
        NAT.check(NAT.lib.ddrs_mc_window_fwd_gauged(ws.handle, cp, T, ptr(qp), ptr(q0), ptr(Qg),
                                                    NAT.FORM_PERSISTENT, stream))
        NAT.check(NAT.lib.ddrs_mc_window_adj_gauged(ws.handle, cp, T, ptr(qp), ptr(q0),
                                                    ptr(gQg), ptr(gn), ptr(gq), None, None, None,
                                                    NAT.FORM_PERSISTENT, stream))

    trials = []
    for i in range(3):
        qpi = NW.gen_forcing_device(R, T, 94 + i, True, device=DEV)
        q0i = torch.empty(R, device=DEV)
        hot = net.workspace()
        NAT.check(NAT.lib.ddrs_hotstart(hot.handle, cp, ptr(qpi[0]), ptr(q0i), s0))
        gi = torch.rand((T, K), device=DEV, generator=torch.Generator(DEV).manual_seed(i))
        trials.append((qpi, q0i, gi))
    # eager results
    want = []
    for qpi, q0i, gi in trials:
        qp.copy_(qpi)
        q0.copy_(q0i)
        gQg.copy_(gi)
        pair(torch.cuda.current_stream().cuda_stream)
        ws.check()
        want.append((Qg.clone(), gn.clone(), gq.clone()))
    # capture (on torch's capture stream), then replay each trial twice
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        pair(torch.cuda.current_stream().cuda_stream)
    for rep in range(2):
        for i, (qpi, q0i, gi) in enumerate(trials):
            qp.copy_(qpi)
            q0.copy_(q0i)
            gQg.copy_(gi)
            Qg.fill_(-1.0)
            gn.zero_()
            gq.zero_()
            graph.replay()
            torch.cuda.synchronize()
            ws.check()
            say(f"replay {rep} trial {i}")
            for got, w, name in zip((Qg, gn, gq), want[i], ("Qg", "dL/dn", "dL/dq_spatial")):
                assert torch.equal(got, w), f"replay {rep} trial {i}: {name}"
            # eager work between replays
            _ = (qpi * 2.0).sum().item()
    # the host cannot know which forward the graph left in the history: an
    # eager adjoint now needs an eager forward first
    with pytest.raises(NAT.DdrsError):
        NAT.check(NAT.lib.ddrs_mc_window_adj_gauged(ws.handle, cp, T, ptr(qp), ptr(q0),
                                                    ptr(gQg), ptr(gn), ptr(gq), None, None, None,
                                                    NAT.FORM_PERSISTENT,
                                                    torch.cuda.current_stream().cuda_stream))
    qpi, q0i, gi = trials[1]
    qp.copy_(qpi)
    q0.copy_(q0i)
    gQg.copy_(gi)
    pair(torch.cuda.current_stream().cuda_stream)
    ws.check()
    for got, w in zip((Qg, gn, gq), want[1]):
        assert torch.equal(got, w)

"""Autotuning of a routing workload's launch layout: the device schedule's
longest chain of links per block (sub-level cap) and block size, timed on
short probe windows of the workload's own network, and the checkpoint window,
the longest whose workspace fits the memory budget (no recompute when the
whole window fits). Results do not depend on any of these choices; only time
and memory do.

Model. A persistent window is a pipeline over the schedule's chunks: its time
is a fill term (chunk hops from the headwaters to the outlet) plus one
per-step term per timestep (the slowest chunk's cycle). A lower cap shortens
each chunk's solve but adds hops, so the best cap depends on the window
length. Each candidate is timed at two probe lengths and the line through
them is extrapolated to the workload's length; when the workload is short,
it is timed at its own length instead.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .config import Params
from .routing import hotstart
from .sparse import Network, SparseAdjacency, Workspace, _ptr, _stream


@dataclass
class TunedSchedule:
    """The chosen layout and what the probes measured.

    network: the plan built with the chosen options (pass it on);
    window: checkpoint window in timesteps (T - 1: one segment);
    predicted_ms: forward + adjoint of the whole window, extrapolated;
    probes: one row per candidate (sub_levels, block, predicted_ms, fwd_ms_per_step,
    total_ms_per_step)."""

    max_sub_levels: int
    block: int
    window: int
    predicted_ms: float
    network: Network
    probes: List[dict] = field(default_factory=list)


def choose_window(net: Network, T: int, budget_bytes: int, dense_grad: bool = False) -> int:
    """The longest checkpoint window (at most T - 1) whose workspace fits
    budget_bytes. The size is not monotone in the window (the window's state
    grows with it, the checkpoints shrink), so windows are tried from the
    longest down (each try is a host-side size query). Raises MemoryError when
    none fits."""
    steps = max(int(T) - 1, 1)
    smallest = None
    for w in range(steps, 0, -1):
        b = net.workspace_bytes(w, steps, dense_grad)
        if b <= budget_bytes:
            return w
        smallest = b if smallest is None else min(smallest, b)
    raise MemoryError(f"a workspace for {steps} steps needs at least {smallest} bytes at any "
                      f"checkpoint window; the budget is {budget_bytes}")


def _default_budget() -> int:
    free, _ = torch.cuda.mem_get_info()
    return int(0.9 * free)


def _probe(net: Network, params: Params, adj: SparseAdjacency, lengths: Sequence[int],
           window: int, q_prime: Optional[torch.Tensor], n: Optional[torch.Tensor],
           q_spatial: Optional[torch.Tensor], gauges, repeats: int, dev) -> list:
    """(forward ms, forward + adjoint ms) of persistent gauged windows of each
    length (timesteps) with checkpoint window at most `window`, minimum over
    `repeats` runs after one warm-up run."""
    R = adj.n
    steps = max(lengths)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    nn = n if n is not None else torch.full((R,), 0.05, device=dev)
    qs = q_spatial if q_spatial is not None else torch.full((R,), 0.5, device=dev)
    p = torch.full((R,), float(params.defaults["p_spatial"]), device=dev)
    length = f(adj.length_m)
    slope = torch.clamp_min(f(adj.slope), params.attribute_minimums.slope)
    if q_prime is not None and q_prime.shape[0] >= steps + 1:
        qp = q_prime[: steps + 1].contiguous()
    else:
        qp = torch.ones((steps + 1, R), device=dev)
    q0 = hotstart(net, params, qp[0])
    cp = params.to_c()
    ws = Workspace(net, min(steps, int(window)), steps, dense_grad=False)
    try:
        ws.set_gauges(np.asarray(gauges if gauges is not None else [R - 1], np.int32))
        s = _stream()
        N.check(N.lib.ddrs_mc_bind_statics(ws.handle, cp, _ptr(nn), _ptr(qs), _ptr(p),
                                           _ptr(length), _ptr(slope), None, s))
        K = ws.gauges.size
        Qg = torch.empty((steps + 1, K), device=dev)
        gQg = torch.full((steps + 1, K), 1e-3, device=dev)
        gn, gq = torch.empty(R, device=dev), torch.empty(R, device=dev)
        stream = torch.cuda.current_stream()

        def run(T):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record(stream)
            N.check(N.lib.ddrs_mc_window_fwd_gauged(ws.handle, cp, T, _ptr(qp), _ptr(q0),
                                                    _ptr(Qg), N.FORM_PERSISTENT, s))
            e[1].record(stream)
            N.check(N.lib.ddrs_mc_window_adj_gauged(ws.handle, cp, T, _ptr(qp), _ptr(q0),
                                                    _ptr(gQg), _ptr(gn), _ptr(gq), None, None,
                                                    None, N.FORM_PERSISTENT, s))
            e[2].record(stream)
            ws.check()
            return e[0].elapsed_time(e[1]), e[0].elapsed_time(e[2])

        run(lengths[0] + 1)
        out = []
        for L in lengths:
            best = min((run(L + 1) for _ in range(repeats)), key=lambda v: v[1])
            out.append(best)
        return out
    finally:
        ws.close()


def tune_schedule(adj: SparseAdjacency, T: int, params: Optional[Params] = None,
                  q_prime: Optional[torch.Tensor] = None, n: Optional[torch.Tensor] = None,
                  q_spatial: Optional[torch.Tensor] = None, gauges=None,
                  sub_levels: Sequence[int] = (2, 4, 8), blocks: Sequence[int] = (64,),
                  probe_steps: Sequence[int] = (96, 288), budget_bytes: Optional[int] = None,
                  dense_grad: bool = False, repeats: int = 2,
                  device="cuda") -> TunedSchedule:
    """Pick the sub-level cap and block of the device schedule, and the
    checkpoint window, for forward + adjoint windows of T timesteps on `adj`.

    q_prime [>= probe + 1, n], n and q_spatial (device tensors) are the
    workload's own inputs for the probes when given (constants otherwise);
    gauges: the readout reaches (the last reach otherwise). The window is the
    longest that fits budget_bytes (default 90 % of the free device memory) at
    the chosen layout; the prediction adds the recomputed forward segments
    when it is shorter than T - 1."""
    params = params if params is not None else Params()
    dev = torch.device(device)
    steps = max(int(T) - 1, 1)
    budget = int(budget_bytes) if budget_bytes is not None else _default_budget()
    direct = steps <= int(2.5 * max(probe_steps))
    lengths = [steps] if direct else sorted(int(x) for x in probe_steps)
    probes, best = [], None
    for blk in blocks:
        for sl in sub_levels:
            net = Network(adj, int(sl), int(blk))
            window = choose_window(net, T, budget, dense_grad)
            r = _probe(net, params, adj, lengths, window, q_prime, n, q_spatial, gauges,
                       repeats, dev)
            if direct:
                # timed as it will run, recomputed segments included
                f_step = r[0][0] / steps
                tot = r[0][1]
                step_ms = tot / steps
                pred = tot
            else:
                (f1, t1), (f2, t2) = r[0], r[-1]
                L1, L2 = lengths[0], lengths[-1]
                step_ms = (t2 - t1) / (L2 - L1)
                f_step = (f2 - f1) / (L2 - L1)
                tot = t2 + step_ms * (steps - L2)
                last = steps - window * ((steps - 1) // window)   # length of the last segment
                pred = tot + f_step * (steps - last)                # recomputed forward segments
            probes.append(dict(sub_levels=int(sl), block=int(blk), window=window,
                               predicted_ms=pred, fwd_ms_per_step=f_step,
                               total_ms_per_step=step_ms, probe_lengths=lengths))
            if best is None or pred < best[0]:
                best = (pred, sl, blk, window, net)
            else:
                del net
    pred, sl, blk, window, net = best
    return TunedSchedule(int(sl), int(blk), int(window), float(pred), net, probes)

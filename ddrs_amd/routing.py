"""Host-side mirror of the reference routing interface on the HIP library.

denormalize        src/routing/utils.rs:32-41
timestep_forward   src/routing/mmc_op.rs:1704-1857 (TimestepOp, parents
                   [n, q_spatial, p_spatial, q_t, q'_t]; backward 850-893)
MuskingumCunge     src/routing/mmc.rs:189-632 (setup_inputs, forward); the T
                   timestep nodes of the reference tape become one window
                   forward and one reverse-time adjoint on the GPU.

Everything runs on the GPU through ddrs_amd/libddrs_hip.so; there is no CPU
path. Tensors must be CUDA (HIP) float32 tensors.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N
from .config import Params
from .sparse import Network, SparseAdjacency, _dev_f32, _ptr, _stream

_tokens = itertools.count(1)


def denormalize(value: torch.Tensor, bounds, log_space: bool) -> torch.Tensor:
    """[0, 1] network output to physical units (routing/utils.rs:32-41)."""
    lo, hi = float(bounds[0]), float(bounds[1])
    if log_space:
        log_min = math.log(lo + 1e-6)
        log_max = math.log(hi)
        return torch.exp(value * (log_max - log_min) + log_min)
    return value * (hi - lo) + lo


def _bind(ws, params: Params, n, qsp, p, length, slope, xst):
    cp = params.to_c()
    R = ws.net.n
    args = [_dev_f32(v, R) for v in (n, qsp, p, length, slope)]
    x = _dev_f32(xst, R) if xst is not None else None
    N.check(N.lib.ddrs_mc_bind_statics(ws.handle, cp, *[_ptr(a) for a in args], _ptr(x),
                                       _stream()), "ddrs_mc_bind_statics")
    return cp, args, x


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, qsp, p, q_t, qp_t, net, params, length, slope, xst):
        ws = net.workspace()
        cp, keep, xk = _bind(ws, params, n, qsp, p, length, slope, xst)
        q = _dev_f32(q_t, net.n)
        qp = _dev_f32(qp_t, net.n)
        q_next = torch.empty_like(q)
        x_sol = torch.empty_like(q)
        N.check(N.lib.ddrs_mc_step_fwd(ws.handle, cp, _ptr(q), _ptr(qp), _ptr(q_next),
                                       _ptr(x_sol), _stream()), "ddrs_mc_step_fwd")
        ws.check()
        ctx.net, ctx.params = net, params
        ctx.save_for_backward(*keep, q, qp, x_sol, *( [xk] if xk is not None else []))
        ctx.has_x = xk is not None
        return q_next

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        n, qsp, p, length, slope, q, qp, x_sol = saved[:8]
        xst = saved[8] if ctx.has_x else None
        net = ctx.net
        ws = net.workspace()
        cp, _, _ = _bind(ws, ctx.params, n, qsp, p, length, slope, xst)
        gy = _dev_f32(g, net.n)
        outs = [torch.empty_like(gy) for _ in range(5)]
        N.check(N.lib.ddrs_mc_step_bwd(ws.handle, cp, _ptr(q), _ptr(qp), _ptr(x_sol), _ptr(gy),
                                       *[_ptr(o) for o in outs], _stream()), "ddrs_mc_step_bwd")
        ws.check()
        gn, gqsp, gp, gq, gqp = outs
        return gn, gqsp, gp, gq, gqp, None, None, None, None, None


def timestep_forward(params: Params, net: Network, n, q_spatial, p_spatial, q_t, q_prime_t,
                     length, slope, x_storage=None) -> torch.Tensor:
    """One fused Muskingum-Cunge step (TimestepOp). q_prime_t must already be
    clamped to the discharge floor; slope already clamped to the slope floor."""
    return _Step.apply(n, q_spatial, p_spatial, q_t, q_prime_t, net, params, length, slope,
                       x_storage)


class _Hotstart(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qp0, net, params):
        ws = net.workspace()
        cp = params.to_c()
        b = _dev_f32(qp0, net.n)
        q0 = torch.empty_like(b)
        N.check(N.lib.ddrs_hotstart(ws.handle, cp, _ptr(b), _ptr(q0), _stream()), "ddrs_hotstart")
        ws.check()
        ctx.net, ctx.params = net, params
        ctx.save_for_backward(q0)
        return q0

    @staticmethod
    def backward(ctx, g):
        (q0,) = ctx.saved_tensors
        net = ctx.net
        ws = net.workspace()
        gg = _dev_f32(g, net.n)
        out = torch.empty_like(gg)
        N.check(N.lib.ddrs_hotstart_adj(ws.handle, ctx.params.to_c(), _ptr(q0), _ptr(gg),
                                        _ptr(out), _stream()), "ddrs_hotstart_adj")
        ws.check()
        return out, None, None


def hotstart(net: Network, params: Params, q_prime0: torch.Tensor) -> torch.Tensor:
    """Cold start Q0 = max(solve(I - N, q'_0), discharge_lb) (mmc.rs:326-361)."""
    return _Hotstart.apply(q_prime0, net, params)


class _Window(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, qsp, p, q_prime, q0, net, params, length, slope, xst, form, window):
        T = int(q_prime.shape[0])
        ws = net.workspace(window, max(T - 1, 1))
        cp, keep, xk = _bind(ws, params, n, qsp, p, length, slope, xst)
        qp = _dev_f32(q_prime, T * net.n)
        q0d = _dev_f32(q0, net.n)
        Q = torch.empty((T, net.n), dtype=torch.float32, device=qp.device)
        N.check(N.lib.ddrs_mc_window_fwd(ws.handle, cp, T, _ptr(qp), _ptr(q0d), _ptr(Q), form,
                                         _stream()), "ddrs_mc_window_fwd")
        ws.check()
        token = next(_tokens)
        ws.fwd_token = token
        ctx.token, ctx.net, ctx.params, ctx.form, ctx.window, ctx.T = token, net, params, form, window, T
        ctx.has_x = xk is not None
        ctx.save_for_backward(*keep, qp, q0d, *([xk] if xk is not None else []))
        return Q

    @staticmethod
    def backward(ctx, gQ):
        saved = ctx.saved_tensors
        n, qsp, p, length, slope, qp, q0d = saved[:7]
        xst = saved[7] if ctx.has_x else None
        net, T = ctx.net, ctx.T
        ws = net.workspace(ctx.window, max(T - 1, 1))
        cp, _, _ = _bind(ws, ctx.params, n, qsp, p, length, slope, xst)
        if ws.fwd_token != ctx.token:
            N.check(N.lib.ddrs_mc_window_fwd(ws.handle, cp, T, _ptr(qp), _ptr(q0d), None,
                                             N.FORM_PERSISTENT, _stream()), "ddrs_mc_window_fwd")
            ws.fwd_token = ctx.token
        g = _dev_f32(gQ, T * net.n)
        R = net.n
        gn = torch.empty(R, dtype=torch.float32, device=g.device)
        gqsp, gp, gq0 = torch.empty_like(gn), torch.empty_like(gn), torch.empty_like(gn)
        gqp = torch.empty((T, R), dtype=torch.float32, device=g.device)
        N.check(N.lib.ddrs_mc_window_adj(ws.handle, cp, T, _ptr(qp), _ptr(q0d), _ptr(g), _ptr(gn),
                                         _ptr(gqsp), _ptr(gp), _ptr(gqp), _ptr(gq0), ctx.form,
                                         _stream()), "ddrs_mc_window_adj")
        ws.check()
        return gn, gqsp, gp, gqp, gq0, None, None, None, None, None, None, None


def route_window(net: Network, params: Params, n, q_spatial, p_spatial, q_prime, q0, length,
                 slope, x_storage=None, form: int = N.FORM_PERSISTENT,
                 window: Optional[int] = None) -> torch.Tensor:
    """Q [T, R] of the reference time loop (mmc.rs:519-632) from physical
    parameters, raw q' [T, R] and initial state q0 [R]; differentiable in
    n, q_spatial, p_spatial, q' and q0."""
    T = int(q_prime.shape[0])
    window = max(T - 1, 1) if window is None else int(window)
    return _Window.apply(n, q_spatial, p_spatial, q_prime, q0, net, params, length, slope,
                         x_storage, form, window)


class _WindowGauged(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n, qsp, p, q_prime, q0, net, params, length, slope, xst, form, window,
                gauges, q_sched):
        T = int(q_prime.shape[0])
        sched = q_sched is not None
        ws = _gauge_workspace(net, window, max(T - 1, 1), gauges, sched)
        cp, keep, xk = _bind(ws, params, n, qsp, p, length, slope, xst)
        # the rows the kernels read: q' itself, or its schedule-order layout
        qp = _dev_f32(q_sched, T * net.positions) if sched else _dev_f32(q_prime, T * net.n)
        q0d = _dev_f32(q0, net.n)
        K = int(ws.gauges.size)
        Qg = torch.empty((T, K), dtype=torch.float32, device=qp.device)
        N.check(N.lib.ddrs_mc_window_fwd_gauged(ws.handle, cp, T, _ptr(qp), _ptr(q0d), _ptr(Qg),
                                                form, _stream()), "ddrs_mc_window_fwd_gauged")
        ws.check()
        token = next(_tokens)
        ws.fwd_token = token
        ctx.token, ctx.net, ctx.params, ctx.form, ctx.window, ctx.T = token, net, params, form, window, T
        ctx.gauges, ctx.sched = ws.gauges, sched
        ctx.has_x = xk is not None
        ctx.save_for_backward(*keep, qp, q0d, *([xk] if xk is not None else []))
        return Qg

    @staticmethod
    def backward(ctx, gQg):
        saved = ctx.saved_tensors
        n, qsp, p, length, slope, qp, q0d = saved[:7]
        xst = saved[7] if ctx.has_x else None
        net, T = ctx.net, ctx.T
        ws = _gauge_workspace(net, ctx.window, max(T - 1, 1), ctx.gauges, ctx.sched)
        cp, _, _ = _bind(ws, ctx.params, n, qsp, p, length, slope, xst)
        if ws.fwd_token != ctx.token:
            N.check(N.lib.ddrs_mc_window_fwd(ws.handle, cp, T, _ptr(qp), _ptr(q0d), None,
                                             N.FORM_PERSISTENT, _stream()), "ddrs_mc_window_fwd")
            ws.fwd_token = ctx.token
        g = _dev_f32(gQg, T * int(ctx.gauges.size))
        R = net.n
        gn = torch.empty(R, dtype=torch.float32, device=g.device)
        gqsp, gp = torch.empty_like(gn), torch.empty_like(gn)
        need_qp, need_q0 = ctx.needs_input_grad[3], ctx.needs_input_grad[4]
        gqp = torch.empty((T, R), dtype=torch.float32, device=g.device) if need_qp else None
        gq0 = torch.empty_like(gn) if need_q0 else None
        N.check(N.lib.ddrs_mc_window_adj_gauged(ws.handle, cp, T, _ptr(qp), _ptr(q0d), _ptr(g),
                                                _ptr(gn), _ptr(gqsp), _ptr(gp), _ptr(gqp),
                                                _ptr(gq0), ctx.form, _stream()),
                "ddrs_mc_window_adj_gauged")
        ws.check()
        return gn, gqsp, gp, gqp, gq0, None, None, None, None, None, None, None, None, None


def _gauge_workspace(net: Network, window: int, max_steps: int, gauges,
                     schedule_forcing: bool = False) -> "Workspace":
    """A workspace of `net` without the dense dL/dQ buffer, its gauges set to
    `gauges` (kept while unchanged), taking q' in reach or (schedule_forcing)
    schedule order. One that holds at least `window` and `max_steps` is reused
    (a longer checkpoint window gives the same results); when none does, the
    smaller ones are freed before the new one is made, so that windows of
    growing length do not pile up device memory. A backward whose forward ran
    on a freed workspace recomputes its forward."""
    from .sparse import Workspace
    window, max_steps, sf = int(window), int(max_steps), bool(schedule_forcing)
    gauged = [k for k in net._ws if k[0] == "gauged"]
    ws = next((net._ws[k] for k in gauged
               if k[1] >= window and k[2] >= max_steps and k[3] == sf), None)
    if ws is None:
        for k in gauged:
            net._ws.pop(k).close()
        ws = Workspace(net, window, max_steps, dense_grad=False, schedule_forcing=sf)
        net._ws[("gauged", window, max_steps, sf)] = ws
    g = np.ascontiguousarray(gauges, np.int32)
    if getattr(ws, "gauges", None) is None or not np.array_equal(ws.gauges, g):
        ws.set_gauges(g)
    return ws


def route_window_gauged(net: Network, params: Params, n, q_spatial, p_spatial, q_prime, q0,
                        length, slope, gauges, x_storage=None, form: int = N.FORM_PERSISTENT,
                        window: Optional[int] = None,
                        q_sched: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Q at the gauge reaches only, [T, K] with K = len(gauges) (reference
    reach ids): the same values as route_window(...)[:, gauges], computed
    without writing the full [T, R] discharge; its backward gives the same
    gradients as route_window's for a loss that reads Q only at the gauges
    (dL/dq' and dL/dq0 only when those inputs require gradients). q_sched:
    q' already in schedule order (net.permute_forcing(q_prime), laid out once
    per batch); the window then reads it in place and permutes nothing, with
    bit-identical results."""
    T = int(q_prime.shape[0])
    window = max(T - 1, 1) if window is None else int(window)
    if q_sched is not None and tuple(q_sched.shape) != (T, net.positions):
        raise ValueError(f"q_sched must be [{T}, {net.positions}], got {tuple(q_sched.shape)}")
    return _WindowGauged.apply(n, q_spatial, p_spatial, q_prime, q0, net, params, length, slope,
                               x_storage, form, window, gauges, q_sched)


@dataclass
class RoutingInputs:
    adjacency: SparseAdjacency
    x_storage: Optional[torch.Tensor] = None


@dataclass
class SpatialParameters:
    """Network outputs in [0, 1]; denormalized inside setup_inputs."""

    n: torch.Tensor
    q_spatial: torch.Tensor
    p_spatial: Optional[torch.Tensor] = None


class MuskingumCunge:
    """Differentiable Muskingum-Cunge routing engine (mmc.rs:189-632)."""

    def __init__(self, params: Params, device="cuda", form: int = N.FORM_PERSISTENT,
                 window: Optional[int] = None):
        self.params = params
        self.device = torch.device(device)
        self.form = form
        self.window = window
        self.net: Optional[Network] = None
        self._adj = None
        self.discharge_t: Optional[torch.Tensor] = None

    def setup_inputs(self, inputs: RoutingInputs, streamflow: torch.Tensor,
                     params: SpatialParameters, carry_state: bool = False,
                     initial_state: Optional[torch.Tensor] = None) -> None:
        adj = inputs.adjacency
        if self.net is None or self._adj is not adj:
            self.net = Network(adj)
            self._adj = adj
        R = adj.n
        dev = self.device
        P = self.params
        self.length = torch.as_tensor(adj.length_m, dtype=torch.float32, device=dev)
        self.slope = torch.clamp_min(torch.as_tensor(adj.slope, dtype=torch.float32, device=dev),
                                     P.attribute_minimums.slope)
        self.x_storage = inputs.x_storage
        self.q_prime = streamflow
        rng = P.parameter_ranges
        log = set(P.log_space_parameters)
        self.n = denormalize(params.n, rng.n, "n" in log)
        self.q_spatial = denormalize(params.q_spatial, rng.q_spatial, "q_spatial" in log)
        if params.p_spatial is not None:
            self.p_spatial = denormalize(params.p_spatial, rng.p_spatial, "p_spatial" in log)
        else:
            self.p_spatial = torch.full((R,), float(P.defaults["p_spatial"]), dtype=torch.float32,
                                        device=dev)
        lb = P.attribute_minimums.discharge
        if initial_state is not None:
            self.discharge_t = torch.clamp_min(initial_state, lb)
        elif not carry_state or self.discharge_t is None:
            self.discharge_t = hotstart(self.net, P, streamflow[0])

    def forward(self) -> torch.Tensor:
        """Routed discharge [n, T] (a transposed view of the time-major result)."""
        Q = route_window(self.net, self.params, self.n, self.q_spatial, self.p_spatial,
                         self.q_prime, self.discharge_t, self.length, self.slope, self.x_storage,
                         form=self.form, window=self.window)
        self.discharge_t = Q[-1]
        return Q.t()

    def discharge_state(self) -> Optional[torch.Tensor]:
        return self.discharge_t

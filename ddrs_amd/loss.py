"""Gauge readout and training loss on the GPU (ddrs_amd/csrc/ddrs_loss.hip).

Host mirror of the reference's training-side functions, same names and
argument meaning:
- scatter_add_by_group (src/training/forward.rs:18-34);
- tau_trim_and_downsample (src/training/loss.rs:33-95);
- nse_batch_loss (src/training/loss.rs:226-237);
- gauge_nse_loss: the micro-batch objective of src/training/driver.rs:110-201
  (readout, pooling, warmup slice, NaN-gauge filter, batch NSE) in one call,
  with its gradient straight into the routed discharge.
All run in the HIP kernels through the C ABI; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from .sparse import _ptr, _stream


def _cuda_f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    return t.detach().to(torch.float32).contiguous()


class GaugePlan:
    """The outflow map of a batch (RoutingTensors::flat_indices / group_ids,
    src/data/dataset.rs:139-141): entry k adds reach flat_indices[k] into gauge
    group_ids[k]. Holds the device CSR and scratch for series of up to
    max_t_hours hours."""

    def __init__(self, n_reaches: int, flat_indices: Sequence[int], group_ids: Sequence[int],
                 num_gauges: int, max_t_hours: int):
        flat = np.ascontiguousarray(flat_indices, np.int32)
        grp = np.ascontiguousarray(group_ids, np.int32)
        if flat.shape != grp.shape or flat.ndim != 1:
            raise ValueError("flat_indices and group_ids must be 1-D of equal length")
        h = N.P()
        N.check(N.lib.ddrs_gauge_plan_create(int(n_reaches), int(num_gauges), flat.size,
                                             flat.ctypes.data, grp.ctypes.data, int(max_t_hours),
                                             ctypes.byref(h)), "ddrs_gauge_plan_create")
        self.handle = h
        self.n_reaches, self.num_gauges, self.max_t_hours = int(n_reaches), int(num_gauges), int(max_t_hours)

    @classmethod
    def from_outflow_idx(cls, n_reaches: int, outflow_idx: Sequence[Sequence[int]],
                         max_t_hours: int) -> "GaugePlan":
        """flat / group from per-gauge outflow reach lists, as
        RoutingBatch::to_tensors builds them (src/data/dataset.rs:160-168)."""
        flat = [int(s) for segs in outflow_idx for s in segs]
        grp = [g for g, segs in enumerate(outflow_idx) for _ in segs]
        return cls(n_reaches, flat, grp, len(outflow_idx), max_t_hours)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and N is not None and getattr(N, "lib", None) is not None:
            N.lib.ddrs_gauge_plan_destroy(h)
        self.handle = None


def _strides(runoff: torch.Tensor, time_major: bool) -> Tuple[int, int, int]:
    """(T, stride_n, stride_t) of a contiguous [N, T] or time-major [T, N] series."""
    if time_major:
        T, n = runoff.shape
        return T, 1, n
    n, T = runoff.shape
    return T, T, 1


class _Scatter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, runoff, plan, time_major):
        x = _cuda_f32(runoff, "runoff")
        T, sn, st = _strides(x, time_major)
        out = torch.empty(plan.num_gauges, T, device=x.device, dtype=torch.float32)
        N.check(N.lib.ddrs_scatter_add_by_group(plan.handle, _ptr(x), T, sn, st, _ptr(out),
                                                _stream()), "ddrs_scatter_add_by_group")
        ctx.plan, ctx.time_major, ctx.shape = plan, time_major, x.shape
        return out

    @staticmethod
    def backward(ctx, g):
        g = g.detach().to(torch.float32).contiguous()
        gr = torch.zeros(ctx.shape, device=g.device, dtype=torch.float32)
        T, sn, st = _strides(gr, ctx.time_major)
        N.check(N.lib.ddrs_scatter_add_by_group_bwd(ctx.plan.handle, _ptr(g), T, sn, st, _ptr(gr),
                                                    _stream()), "ddrs_scatter_add_by_group_bwd")
        return gr, None, None


def scatter_add_by_group(runoff: torch.Tensor, plan: GaugePlan,
                         time_major: bool = False) -> torch.Tensor:
    """[N, T] (or [T, N] with time_major) -> [G, T]: per gauge, the sum of its
    outflow reaches (src/training/forward.rs:18-34)."""
    return _Scatter.apply(runoff, plan, time_major)


class _Pool(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, tau):
        x = _cuda_f32(x, "predictions_hourly")
        G, T = x.shape
        out = torch.empty(G, (T - 24) // 24 if T >= 48 else 0, device=x.device)
        N.check(N.lib.ddrs_tau_trim_downsample(_ptr(x), G, T, int(tau), _ptr(out), _stream()),
                "ddrs_tau_trim_downsample")
        ctx.tau, ctx.shape = int(tau), x.shape
        return out

    @staticmethod
    def backward(ctx, g):
        g = g.detach().to(torch.float32).contiguous()
        G, T = ctx.shape
        gin = torch.empty(G, T, device=g.device)
        N.check(N.lib.ddrs_tau_trim_downsample_bwd(_ptr(g), G, T, ctx.tau, _ptr(gin), _stream()),
                "ddrs_tau_trim_downsample_bwd")
        return gin, None


def tau_trim_and_downsample(predictions_hourly: torch.Tensor, tau: int) -> torch.Tensor:
    """[G, T_hours] -> [G, (T_hours - 24) // 24] (src/training/loss.rs:33-95)."""
    return _Pool.apply(predictions_hourly, tau)


class _Nse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, o, sigma, eps, plan):
        p, o, s = _cuda_f32(p, "p"), _cuda_f32(o, "o"), _cuda_f32(sigma, "sigma")
        G, T = p.shape
        loss = torch.empty((), device=p.device)
        N.check(N.lib.ddrs_nse_batch_loss(plan.handle, _ptr(p), _ptr(o), _ptr(s), G, T,
                                          float(eps), _ptr(loss), _stream()),
                "ddrs_nse_batch_loss")
        ctx.save_for_backward(p, o, s)
        ctx.eps = float(eps)
        return loss

    @staticmethod
    def backward(ctx, gl):
        p, o, s = ctx.saved_tensors
        gl = gl.detach().to(torch.float32).contiguous()
        gp = torch.empty_like(p)
        G, T = p.shape
        N.check(N.lib.ddrs_nse_batch_loss_bwd(_ptr(p), _ptr(o), _ptr(s), G, T, ctx.eps, _ptr(gl),
                                              _ptr(gp), _stream()), "ddrs_nse_batch_loss_bwd")
        return gp, None, None, None, None


def nse_batch_loss(p: torch.Tensor, o: torch.Tensor, sigma: torch.Tensor, eps: float,
                   plan: GaugePlan) -> torch.Tensor:
    """Mean over (gauge, day) of (p - o)^2 / (sigma + eps)^2
    (src/training/loss.rs:226-237); `plan` lends its scratch (G <= its gauges)."""
    return _Nse.apply(p, o, sigma, eps, plan)


_KINDS = {"l1": N.LOSS_L1, "nse-batch": N.LOSS_NSE_BATCH, "nnse-kge": N.LOSS_NNSE_KGE,
          "kge": N.LOSS_KGE}
# LossConfig's defaults (src/config.rs:250-262)
LOSS_WEIGHTS = dict(nnse_weight=1.0, kge_weight=1.0, r_weight=1.0, alpha_weight=1.0,
                    beta_weight=1.0, kge_clamp=10.0)


def _loss_config(kind: str, eps: float, weights: dict) -> "N.LossConfig":
    w = {**LOSS_WEIGHTS, **(weights or {})}
    unknown = set(w) - set(LOSS_WEIGHTS)
    if unknown:
        raise ValueError(f"unknown loss weights {sorted(unknown)}")
    return N.LossConfig(_KINDS[kind], float(eps), *[float(w[k]) for k in LOSS_WEIGHTS])


class _GaugeNse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, runoff, plan, time_major, tau, warmup, obs, sigma, eps, kind, weights):
        x = _cuda_f32(runoff, "runoff")
        o = _cuda_f32(obs, "obs")
        s = _cuda_f32(sigma, "sigma") if sigma is not None else None
        T, sn, st = _strides(x, time_major)
        loss = torch.empty((), device=x.device)
        n_valid = torch.empty((), device=x.device, dtype=torch.int32)
        hourly = torch.empty(plan.num_gauges, T, device=x.device)
        g = torch.zeros_like(x) if ctx.needs_input_grad[0] else None
        cfg = _loss_config(kind, eps, weights)
        N.check(N.lib.ddrs_gauge_batch_loss(plan.handle, ctypes.byref(cfg), _ptr(x), T, sn, st,
                                            int(tau), int(warmup), _ptr(o), _ptr(s), _ptr(loss),
                                            _ptr(n_valid), _ptr(g), _ptr(hourly), _stream()),
                "ddrs_gauge_batch_loss")
        ctx.save_for_backward(g if g is not None else torch.empty(0, device=x.device))
        ctx.mark_non_differentiable(n_valid, hourly)
        return loss, n_valid, hourly

    @staticmethod
    def backward(ctx, gl, _gn, _gh):
        (g,) = ctx.saved_tensors
        return g * gl.to(torch.float32), None, None, None, None, None, None, None, None, None


def gauge_nse_loss(runoff: torch.Tensor, plan: GaugePlan, tau: int, warmup: int,
                   obs: torch.Tensor, sigma: torch.Tensor, eps: float = 0.1,
                   time_major: bool = True):
    """The micro-batch objective of src/training/driver.rs:110-201 with
    loss.kind nse-batch, from the routed discharge (time-major [T, N] by
    default, the layout of route_window). Returns (loss, n_valid, hourly):
    n_valid is loss_denominator (kept gauges x post-warmup days; 0 and loss 0
    when every gauge has a NaN), hourly the [G, T] readout. obs is
    [G, T_days] with NaN for missing days, sigma [G]."""
    return _GaugeNse.apply(runoff, plan, time_major, tau, warmup, obs, sigma, eps, "nse-batch",
                           None)


def gauge_batch_loss(runoff: torch.Tensor, plan: GaugePlan, tau: int, warmup: int,
                     obs: torch.Tensor, sigma: Optional[torch.Tensor] = None, eps: float = 0.1,
                     kind: str = "l1", time_major: bool = True, **weights):
    """gauge_nse_loss for any loss.kind of batch_loss (src/training/loss.rs:179-373):
    "l1" (the reference's default, mean |p - o|), "nse-batch" (sigma
    required), "nnse-kge" or "kge" (per-gauge objectives averaged over
    gauges; weights nnse_weight, kge_weight, r_weight, alpha_weight,
    beta_weight, kge_clamp as LossConfig, default 1, 1, 1, 1, 1, 10). n_valid
    is loss_denominator: kept gauges x days, or kept gauges for the per-gauge
    kinds."""
    if kind not in _KINDS:
        raise ValueError(f"loss kind must be one of {sorted(_KINDS)}, got {kind!r}")
    if kind == "nse-batch" and sigma is None:
        raise ValueError("loss kind nse-batch needs the per-gauge sigma")
    return _GaugeNse.apply(runoff, plan, time_major, tau, warmup, obs, sigma, eps, kind,
                           weights)

"""One training micro-batch of the reference's driver on the library's
kernels: KAN head -> denormalize -> gauged routing window -> tau trim, daily
pooling and batch NSE -> routing adjoint -> KAN backward -> global-norm
clipping -> AdaDelta or Adam (src/training/driver.rs:93-215 for the
micro-batch, :295-330 for the single-batch optimizer step), with the nse-batch
or L1 objective.

The discharge is read out at the gauges inside the routing window (the
forward writes Q only at the gauge reaches, the adjoint takes dL/dQ only
there), so a batch never holds the [T, N] discharge the reference builds
before its scatter_add_by_group; the loss and every gradient are those of the
reference's dense chain.

AdaDelta (src/training/adadelta.rs:16-23,93-125), Adam
(src/training/optimizer.rs:131-149) and clip_grad_norm
(src/training/optimizer.rs:163-193) update the head's few thousand weights with
torch tensor arithmetic on the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional

import numpy as np
import torch

from . import _native as N
from .config import Params
from .kan import KanHead
from .loss import GaugePlan, gauge_batch_loss
from .routing import denormalize, hotstart, route_window_gauged
from .sparse import Network, SparseAdjacency


class AdaDelta:
    """AdaDelta with the reference's update and defaults (rho 0.9, eps 1e-6,
    both running averages starting at zero), per parameter:

        square_avg = rho * square_avg + (1 - rho) * g^2
        delta      = sqrt(acc_delta + eps) / sqrt(square_avg + eps) * g
        acc_delta  = rho * acc_delta + (1 - rho) * delta^2
        param     -= lr * delta

    The scalars are float32, as the reference's f32 fields are
    (src/training/adadelta.rs:56-75,104-121)."""

    def __init__(self, params: Iterable[torch.nn.Parameter], rho: float = 0.9,
                 epsilon: float = 1e-6):
        self.params: List[torch.nn.Parameter] = list(params)
        self.rho = float(np.float32(rho))
        self.one_minus_rho = float(np.float32(1.0) - np.float32(rho))
        self.epsilon = float(np.float32(epsilon))
        self.state: Dict[int, tuple] = {}

    @torch.no_grad()
    def step(self, lr: float = 1.0) -> None:
        """Update every parameter that has a gradient; others keep their state."""
        lr = float(np.float32(lr))
        for p in self.params:
            g = p.grad
            if g is None:
                continue
            sq, acc = self.state.get(id(p), (None, None))
            if sq is None:
                sq, acc = torch.zeros_like(p), torch.zeros_like(p)
            sq = sq * self.rho + (g * g) * self.one_minus_rho
            delta = torch.sqrt(acc + self.epsilon) / torch.sqrt(sq + self.epsilon) * g
            acc = acc * self.rho + (delta * delta) * self.one_minus_rho
            p.sub_(delta * lr)
            self.state[id(p)] = (sq, acc)


class Adam:
    """Adam with the reference's settings (beta1 0.9, beta2 0.999, eps 1e-8;
    src/training/optimizer.rs:131-149, the default optimizer), the published
    update (Kingma and Ba 2015) that burn's Adam and torch.optim.Adam implement,
    per parameter at step t:

        m = beta1 m + (1 - beta1) g
        v = beta2 v + (1 - beta2) g^2
        param -= lr * (m / (1 - beta1^t)) / (sqrt(v / (1 - beta2^t)) + eps)

    with m and v starting at zero."""

    def __init__(self, params: Iterable[torch.nn.Parameter], beta1: float = 0.9,
                 beta2: float = 0.999, epsilon: float = 1e-8):
        self.params: List[torch.nn.Parameter] = list(params)
        self.beta1, self.beta2, self.epsilon = float(beta1), float(beta2), float(epsilon)
        self.state: Dict[int, tuple] = {}

    @torch.no_grad()
    def step(self, lr: float = 1e-3) -> None:
        for p in self.params:
            g = p.grad
            if g is None:
                continue
            m, v, t = self.state.get(id(p), (None, None, 0))
            if m is None:
                m, v = torch.zeros_like(p), torch.zeros_like(p)
            t += 1
            m = m * self.beta1 + g * (1.0 - self.beta1)
            v = v * self.beta2 + (g * g) * (1.0 - self.beta2)
            m_hat = m / (1.0 - self.beta1 ** t)
            v_hat = v / (1.0 - self.beta2 ** t)
            p.sub_(lr * m_hat / (torch.sqrt(v_hat) + self.epsilon))
            self.state[id(p)] = (m, v, t)

    def moments(self, p: torch.nn.Parameter):
        """(m, v) of a parameter after the last step."""
        m, v, _ = self.state[id(p)]
        return m, v


@torch.no_grad()
def clip_grad_norm(params: Iterable[torch.nn.Parameter], max_norm: float) -> float:
    """Scale every gradient by max_norm / norm when the global L2 norm over all
    gradients exceeds max_norm (no epsilon, as src/training/optimizer.rs:163-193).
    Returns the norm before clipping."""
    grads = [p.grad for p in params if p.grad is not None]
    if not grads:
        return 0.0
    sq = torch.stack([(g.float() * g.float()).sum() for g in grads]).sum()
    norm = float(torch.sqrt(sq))
    if norm <= max_norm or norm == 0.0:
        return norm
    scale = float(np.float32(max_norm) / np.float32(norm))
    for g in grads:
        g.mul_(scale)
    return norm


@dataclass
class GaugedBatch:
    """One micro-batch on the device (the routing side of the reference's
    RoutingTensors, src/data/dataset.rs:120-160, and the observation side of
    run_micro_batch):

    adjacency     the batch's compressed network (reach positions 0..N-1);
    attributes    [N, F] normalised reach attributes, the KAN head's input;
    q_prime       [T, N] hourly lateral inflow (time-major);
    gauges        [G] the gauges' compressed reach positions;
    observations  [G, T_days] daily observations, NaN where missing;
    gauge_obs_std [G] per-gauge standard deviation of the observations (the
                  nse-batch sigma; None for the L1 objective);
    outflow       the batch's outflow_idx (data/collate.rs:297-336) as
                  Compressed.scatter_map() gives it: (reaches read out,
                  flat indices into them, gauge of each entry). None: each
                  gauge's prediction is its own reach.

    The trainer lays q_prime out in the device schedule's order once per
    batch and network (forcing_in_schedule_order) and keeps that copy with the
    batch: a batch whose q_prime is rewritten in place must reset it
    (q_sched = None)."""

    adjacency: SparseAdjacency
    attributes: torch.Tensor
    q_prime: torch.Tensor
    gauges: np.ndarray
    observations: torch.Tensor
    gauge_obs_std: Optional[torch.Tensor]
    outflow: Optional[tuple] = None
    q_sched: Optional[tuple] = field(default=None, repr=False, compare=False)

    def forcing_in_schedule_order(self, net: Network) -> torch.Tensor:
        """q_prime in the schedule order of `net` ([T, net.positions]), made
        on the first call for that network and kept."""
        if self.q_sched is None or self.q_sched[0] is not net:
            self.q_sched = (net, net.permute_forcing(self.q_prime.detach()))
        return self.q_sched[1]

    def readout(self) -> np.ndarray:
        """The reaches whose discharge the routing window reads out."""
        return self.gauges if self.outflow is None else self.outflow[0]


@dataclass
class StepOutcome:
    loss: float
    n_valid: int
    grad_norm: float


class Trainer:
    """The single-batch training path: forward and loss of one micro-batch,
    backward, clip_grad_norm, optimizer step. window: checkpoint segment of the
    routing adjoint in timesteps (None: the whole window in one segment).
    network: a device plan of the batches' adjacency to use (and whose
    workspaces to share) instead of building one."""

    def __init__(self, head: KanHead, params: Optional[Params] = None, tau: int = 9,
                 warmup: int = 5, eps: float = 0.1, grad_clip: float = 1.0,
                 window: Optional[int] = None, form: int = N.FORM_PERSISTENT,
                 network: Optional[Network] = None, loss: str = "nse-batch",
                 optimizer: str = "adadelta", loss_weights: Optional[dict] = None):
        if "n" not in head.learnable or "q_spatial" not in head.learnable:
            raise ValueError("the head must produce n and q_spatial")
        if loss not in ("l1", "nse-batch", "nnse-kge", "kge"):
            raise ValueError(f"loss must be 'l1', 'nse-batch', 'nnse-kge' or 'kge', got {loss!r}")
        if optimizer not in ("adam", "adadelta"):
            raise ValueError(f"optimizer must be 'adam' or 'adadelta', got {optimizer!r}")
        self.head = head
        self.params = params if params is not None else Params()
        self.tau, self.warmup, self.eps, self.grad_clip = int(tau), int(warmup), float(eps), grad_clip
        self.window, self.form = window, form
        self.loss_kind = loss
        self.loss_weights = dict(loss_weights or {})
        self.opt = (AdaDelta if optimizer == "adadelta" else Adam)(head.parameters())
        self._net: Optional[Network] = network
        self._adj = network.adj if network is not None else None
        self._plan: Optional[GaugePlan] = None
        self._plan_key = None
        self._statics = None

    def network(self, batch: GaugedBatch) -> Network:
        if self._net is None or self._adj is not batch.adjacency:
            self._net = Network(batch.adjacency)
            self._adj = batch.adjacency
            self._statics = None
        return self._net

    def _static_inputs(self, batch: GaugedBatch, dev):
        """length, slope floored at its minimum (mmc.rs:259-264), p_spatial at
        its default (used when the head does not produce it), x_storage 0.3
        (training/forward.rs:322-334)."""
        if self._statics is None:
            adj, P = batch.adjacency, self.params
            f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
            length = f(adj.length_m)
            slope = torch.clamp_min(f(adj.slope), P.attribute_minimums.slope)
            p = torch.full((adj.n,), float(P.defaults["p_spatial"]), dtype=torch.float32,
                           device=dev)
            x = torch.full((adj.n,), 0.3, dtype=torch.float32, device=dev)
            self._statics = (length, slope, p, x)
        return self._statics

    def _gauge_plan(self, batch: GaugedBatch, K: int, T: int) -> GaugePlan:
        """scatter_add_by_group of the batch's outflow map over the K reaches
        read out (the identity when each gauge reads its own reach)."""
        if batch.outflow is None:
            key = (K, T, None)
            flat = grp = np.arange(K, dtype=np.int32)
            G = K
        else:
            _, flat, grp = batch.outflow
            G = int(np.asarray(batch.gauges).size)
            key = (K, T, G, hash(np.asarray(flat).tobytes()), hash(np.asarray(grp).tobytes()))
        if self._plan_key != key:
            self._plan = GaugePlan(K, flat, grp, G, T)
            self._plan_key = key
        return self._plan

    def predict(self, batch: GaugedBatch, trace: Optional[dict] = None) -> torch.Tensor:
        """Hourly discharge at the batch's readout reaches (its gauges, or
        their outflow reaches), [T, K], differentiable in the head's
        weights (training/forward.rs:301-387 with the gauge readout). trace: a
        dict that receives the routed physical parameters "n", "q_spatial",
        the initial state "q0" and the gauge discharge "Qg", each keeping its
        gradient after a backward."""
        net = self.network(batch)
        dev = batch.q_prime.device
        length, slope, p, x = self._static_inputs(batch, dev)
        out = self.head(batch.attributes)
        P = self.params
        rng, log = P.parameter_ranges, set(P.log_space_parameters)
        n = denormalize(out["n"], rng.n, "n" in log)
        qsp = denormalize(out["q_spatial"], rng.q_spatial, "q_spatial" in log)
        if "p_spatial" in out:   # learned when the head produces it (forward.rs:319-373)
            p = denormalize(out["p_spatial"], rng.p_spatial, "p_spatial" in log)
        q0 = hotstart(net, P, batch.q_prime[0].detach())
        Qg = route_window_gauged(net, P, n, qsp, p, batch.q_prime.detach(), q0.detach(),
                                 length, slope, batch.readout(), x_storage=x, form=self.form,
                                 window=self.window,
                                 q_sched=batch.forcing_in_schedule_order(net))
        if trace is not None:
            for v in (n, qsp, Qg):
                if v.requires_grad:
                    v.retain_grad()
            trace.update(n=n, q_spatial=qsp, q0=q0, Qg=Qg)
        return Qg

    def loss(self, batch: GaugedBatch, trace: Optional[dict] = None):
        """(loss, n_valid): tau trim and daily pooling, warmup slice, NaN-gauge
        filter and the configured objective (driver.rs:126-201) of
        predict(batch); n_valid is loss_denominator. trace as in predict."""
        Qg = self.predict(batch, trace)
        plan = self._gauge_plan(batch, Qg.shape[1], Qg.shape[0])
        loss, n_valid, _ = gauge_batch_loss(Qg, plan, self.tau, self.warmup, batch.observations,
                                            batch.gauge_obs_std, self.eps, kind=self.loss_kind,
                                            time_major=True, **self.loss_weights)
        return loss, n_valid

    def step(self, batch: GaugedBatch, lr: Optional[float] = None) -> StepOutcome:
        """One optimizer step (driver.rs:315-329) at learning rate lr (default:
        1 for AdaDelta, 1e-3 for Adam). A batch whose gauges are all dropped by
        the NaN filter takes no step (its loss is 0)."""
        if lr is None:
            lr = 1.0 if isinstance(self.opt, AdaDelta) else 1e-3
        for prm in self.head.parameters():
            prm.grad = None
        loss, n_valid = self.loss(batch)
        nv = int(n_valid)
        if nv == 0:
            return StepOutcome(0.0, 0, 0.0)
        loss.backward()
        norm = clip_grad_norm(self.head.parameters(), self.grad_clip)
        self.opt.step(lr)
        return StepOutcome(float(loss.detach()), nv, norm)

    def step_accumulated(self, batches, lr: Optional[float] = None) -> StepOutcome:
        """One optimizer step over a group of micro-batches (the accumulation
        path, driver.rs:360-430): each micro-batch's mean loss is scaled back
        to a sum by its n_valid before its backward, the summed gradients are
        scaled by 1 / sum(n_valid), so that they are the pooled mean's, then
        clipped and stepped. Micro-batches with every gauge dropped add
        nothing; a group of only such takes no step. The outcome's loss is the
        n_valid-weighted mean."""
        if lr is None:
            lr = 1.0 if isinstance(self.opt, AdaDelta) else 1e-3
        return _accumulate_step(self, batches, lr)


def _accumulate_step(tr: "Trainer", batches, lr) -> StepOutcome:
    """Trainer.step_accumulated (a function so that the class stays short)."""
    for prm in tr.head.parameters():
        prm.grad = None
    total_n, weighted = 0, 0.0
    for b in batches:
        loss, n_valid = tr.loss(b)
        n = int(n_valid)
        if n == 0:
            continue
        # the mean back to a sum; the group is renormalised by 1 / total_n
        (loss * float(n)).backward()
        total_n += n
        weighted += float(loss.detach()) * n
    if total_n == 0:
        return StepOutcome(0.0, 0, 0.0)
    scale = float(np.float32(1.0) / np.float32(total_n))
    with torch.no_grad():
        for prm in tr.head.parameters():
            if prm.grad is not None:
                prm.grad.mul_(scale)
    norm = clip_grad_norm(tr.head.parameters(), tr.grad_clip)
    tr.opt.step(lr)
    return StepOutcome(weighted / total_n, total_n, norm)


def upstream_sizes(n: int, rows, cols) -> np.ndarray:
    """Reaches upstream of each reach, itself included, on a tree in
    topological order (each column drains to its row, rows > cols)."""
    parent = np.full(n, -1, np.int64)
    parent[np.asarray(cols, np.int64)] = np.asarray(rows, np.int64)
    size = np.ones(n, np.int64)
    for i in range(n):
        if parent[i] >= 0:
            size[parent[i]] += size[i]
    return size


def nested_gauges(n: int, rows, cols, num_gauges: int, size_lo: float, size_hi: float,
                  seed: int) -> np.ndarray:
    """Gauge reaches whose upstream basins have log-uniform sizes in
    [size_lo, size_hi]: for each drawn size, the unused reach whose upstream
    count is nearest (gauge basins nest, as real gauges along one river do).
    Returned sorted."""
    size = upstream_sizes(n, rows, cols)
    rs = np.random.default_rng(seed)
    targets = np.exp(rs.uniform(math.log(size_lo), math.log(size_hi), num_gauges))
    order = np.argsort(size, kind="stable")
    ss = size[order]
    used = np.zeros(n, bool)
    out = []
    for tg in targets:
        j = int(np.searchsorted(ss, tg))
        lo, hi = j - 1, j
        while True:
            cand = [k for k in (lo, hi) if 0 <= k < n and not used[order[k]]]
            if cand:
                best = min(cand, key=lambda k: (abs(ss[k] - tg), k))
                break
            if lo < 0 and hi >= n:
                raise ValueError("more gauges than reaches")
            lo, hi = lo - 1, hi + 1
        used[order[best]] = True
        out.append(order[best])
    return np.sort(np.asarray(out, np.int32))


def upstream_union(n: int, rows, cols, gauges):
    """The union of the gauges' upstream subgraphs (the edges the reference's
    per-gauge subgraph stores hold, unioned as collate.rs does before
    compression): every edge whose downstream reach lies in some gauge's
    basin. Returns (rows, cols) in the input's order."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    parent = np.full(n, -1, np.int64)
    parent[cols] = rows
    cover = np.zeros(n, bool)
    cover[np.asarray(gauges, np.int64)] = True
    for i in range(n - 1, -1, -1):
        if not cover[i] and parent[i] >= 0 and cover[parent[i]]:
            cover[i] = True
    keep = cover[rows]
    return rows[keep].astype(np.int32), cols[keep].astype(np.int32)

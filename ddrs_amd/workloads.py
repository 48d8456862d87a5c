"""The measured workloads: seeded synthetic river networks, reach attributes,
forcing and gauges of the five configurations, built identically for the
benchmark and the tests.

c1  a 5,000-reach tree, 240 hourly steps, lognormal q'
c2  a 350,000-reach tree (each reach drains to one of the next 256, about 2.7k
    levels), 720 hourly steps, q' with storm pulses, MSE at 64 gauges
c3  512 basins of log-uniform size in [10, 50,000] side by side, 2,160 steps,
    checkpoint segments of 256 steps, the loss at every basin outlet
c4  200,000 basins of Zipf(1.8) size capped at 300,000 and scaled to 2.94 M
    reaches, 8,760 steps, checkpoints every 168 steps, the loss at the outlets
    of the 2,000 largest basins
c5  the training micro-batch: 2,000 nested gauges (basins log-uniform in
    10-50,000 reaches) on a 350,000-reach tree, the union of their basins
    compressed to the batch, 8,760 steps

Seeds: network 1000 + c, attributes 2000 + c, forcing 3000 + c, gauges
4000 + c, basin sizes 7000 + c.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import network as NW
from .sparse import SparseAdjacency

CONFIGS = {
    "c1": dict(R=5000, T=240, window=256, storms=False),
    "c2": dict(R=350000, T=720, window=256, storms=True),
    "c3": dict(basins=512, size_lo=10, size_hi=50000, T=2160, window=256, ckpt=256,
               storms=True),
    "c4": dict(basins=200_000, zipf=1.8, size_hi=300_000, total=2_940_000, T=8760,
               window=256, ckpt=168, storms=True, gauges=2000),
    "c5": dict(R=350_000, window=256, gauges=2000, size_lo=10, size_hi=50_000, T=8760,
               ckpt=None, storms=True, tau=9, warmup=5, eps=0.1, noise=0.1),
}


@dataclass
class Workload:
    """One configuration's network (reference order: every reach drains to a
    larger id), statics and gauges. Forests keep their basins contiguous:
    basin b holds reaches [offsets[b], offsets[b] + sizes[b])."""

    config: str
    R: int
    T: int
    rows: np.ndarray
    cols: np.ndarray
    attrs: dict
    gauges: np.ndarray
    storms: bool
    forcing_seed: int
    sizes: Optional[np.ndarray] = None
    offsets: Optional[np.ndarray] = None
    ckpt: Optional[int] = None
    extra: dict = field(default_factory=dict)

    def adjacency(self) -> SparseAdjacency:
        return SparseAdjacency.from_coo(self.R, self.rows, self.cols, None, self.attrs["length"],
                                        self.attrs["slope"])

    def forcing_device(self, device="cuda:0", T: Optional[int] = None):
        """q' [T, R] generated on the device."""
        return NW.gen_forcing_device(self.R, self.T if T is None else T, self.forcing_seed,
                                     self.storms, device=device)

    def forcing_host(self, T: Optional[int] = None) -> np.ndarray:
        """q' [T, R] generated on the host (the same draws as forcing_device;
        the lognormal and gamma transforms round to fp32 from double)."""
        return NW.gen_forcing(self.R, self.T if T is None else T, self.forcing_seed, self.storms)

    def basin_reaches(self, b: int) -> np.ndarray:
        return np.arange(self.offsets[b], self.offsets[b] + self.sizes[b], dtype=np.int64)

    def sub_basin(self, b: int):
        """Basin b as a network of its own: (rows, cols) relabelled to
        0..sizes[b]-1 (order kept, so its CSR rows keep their column order)."""
        lo, hi = int(self.offsets[b]), int(self.offsets[b] + self.sizes[b])
        keep = (self.cols >= lo) & (self.cols < hi)
        return (self.rows[keep] - lo).astype(np.int32), (self.cols[keep] - lo).astype(np.int32)


def _basin_sizes(cfg: dict, seed: int) -> np.ndarray:
    rs = np.random.default_rng(seed)
    if "zipf" in cfg:
        z = np.minimum(rs.zipf(cfg["zipf"], cfg["basins"]), cfg["size_hi"]).astype(np.float64)
        sizes = np.maximum(1, np.floor(z * cfg["total"] / z.sum())).astype(np.int64)
        sizes[np.argmax(sizes)] += cfg["total"] - int(sizes.sum())
        return sizes
    return np.exp(rs.uniform(np.log(cfg["size_lo"]), np.log(cfg["size_hi"]),
                             cfg["basins"])).astype(np.int64)


def build(config: str) -> Workload:
    """The network, statics and loss gauges of c1..c4."""
    cfg = CONFIGS[config]
    ci = int(config[1:])
    if config == "c5":
        raise ValueError("c5 is a training batch: use build_c5()")
    sizes = offsets = None
    if "basins" in cfg:
        sizes = _basin_sizes(cfg, 7000 + ci)
        rows, cols, outlets = NW.gen_forest(sizes, cfg["window"], 1000 + ci)
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        if cfg.get("gauges"):
            outlets = outlets[np.argsort(sizes)[::-1][: cfg["gauges"]]]
        gauges = np.sort(outlets).astype(np.int32)
        R = int(sizes.sum())
    else:
        R = cfg["R"]
        rows, cols = NW.gen_tree(R, cfg["window"], 1000 + ci)
        rng = np.random.default_rng(4000 + ci)
        gauges = np.sort(np.concatenate([[R - 1], rng.choice(R - 1, 63, replace=False)]))
        gauges = gauges.astype(np.int32)
    attrs = NW.gen_attributes(R, 2000 + ci)
    return Workload(config, R, cfg["T"], rows, cols, attrs, gauges, cfg["storms"], 3000 + ci,
                    sizes, offsets, cfg.get("ckpt"))


def build_c5() -> Workload:
    """The c5 batch: nested gauges on the 350,000-reach tree, the union of their
    basins compressed (positions 0..R-1 in the tree's order). extra holds the
    full tree's size, the active reaches and each gauge's basin size."""
    from . import train as TR
    cfg = CONFIGS["c5"]
    ci = 5
    R0 = cfg["R"]
    rows, cols = NW.gen_tree(R0, cfg["window"], 1000 + ci)
    g_full = TR.nested_gauges(R0, rows, cols, cfg["gauges"], cfg["size_lo"], cfg["size_hi"],
                              seed=7000 + ci)
    basin = TR.upstream_sizes(R0, rows, cols)[g_full]
    ur, uc = TR.upstream_union(R0, rows, cols, g_full)
    active, cr, cc, cg = NW.compress(ur, uc, g_full)
    R = int(active.size)
    at0 = NW.gen_attributes(R0, 2000 + ci)
    attrs = {k: np.ascontiguousarray(v[active]) for k, v in at0.items()}
    return Workload("c5", R, cfg["T"], cr, cc, attrs, cg.astype(np.int32), cfg["storms"],
                    3000 + ci, extra=dict(R0=R0, active=active, basin_sizes=basin,
                                          gauges_full=g_full))


# the KAN parameter head of the batched and training workloads: 10 attributes
# -> 21 hidden -> two KAN blocks (grid 5, order 3) -> n, q_spatial
KAN_SHAPE = dict(F=10, H=21, P=2, L=2, grid=5, k=3)


def kan_head(seed: int, device="cuda:0"):
    """The workloads' KAN head initialised with seed."""
    from . import kan as K
    names = [f"attr_{i}" for i in range(KAN_SHAPE["F"])]
    return K.KanHeadConfig(names, ["n", "q_spatial"], seed, hidden_size=KAN_SHAPE["H"],
                           num_hidden_layers=KAN_SHAPE["L"], grid=KAN_SHAPE["grid"],
                           k=KAN_SHAPE["k"]).init(device)


def c5_batch(wl: Workload, qprime, net, window: Optional[int] = None, device="cuda:0"):
    """The c5 training batch on the device: attributes ~ N(0, 1) from a torch
    generator (seed 5005), and observations from a hidden head's (seed 4242)
    routed discharge at the gauges, tau-trimmed and daily pooled, times
    LogNormal(0, 0.1) noise from the same generator; sigma is each gauge's
    std. The observations do not depend on net's schedule or on window
    (results are bit-identical across both)."""
    import torch
    from . import loss as LS
    from . import train as TR
    cfg = CONFIGS["c5"]
    gen = torch.Generator(device=device).manual_seed(5005)
    x = torch.randn(wl.R, KAN_SHAPE["F"], generator=gen, device=device)
    adj = net.adj
    truth = TR.Trainer(kan_head(4242, device), tau=cfg["tau"], warmup=cfg["warmup"],
                       eps=cfg["eps"], window=window, network=net)
    probe = TR.GaugedBatch(adj, x, qprime, wl.gauges, None, None)
    with torch.no_grad():
        Qg = truth.predict(probe)
        daily = LS.tau_trim_and_downsample(Qg.t().contiguous(), cfg["tau"])
        obs = daily * torch.exp(cfg["noise"] * torch.randn(daily.shape, generator=gen,
                                                           device=device))
        sigma = obs.std(dim=1)
    del truth, Qg, daily
    return TR.GaugedBatch(adj, x, qprime, wl.gauges, obs.contiguous(), sigma.contiguous())


def trees(R: int, rows, cols) -> np.ndarray:
    """The outlet each reach drains to (reference order: every reach drains to
    a larger id, so one pass from the last reach down suffices)."""
    down = np.full(R, -1, np.int64)
    down[np.asarray(cols, np.int64)] = np.asarray(rows, np.int64)
    out = np.arange(R, dtype=np.int64)
    for i in range(R - 1, -1, -1):
        if down[i] >= 0:
            out[i] = out[down[i]]
    return out


class GaugedWindow:
    """The measured step through the C ABI: the statics bound once, the cold
    start on the raw q'[0], the forward window reading discharge out at the
    gauges ([T, K], K = len(gauges)) and the adjoint taking dL/dQ there and
    producing dL/dn and dL/dq_spatial for every reach. The workspace holds no
    dense dL/dQ buffer. window: checkpoint segment (None: the whole window).
    n, q_spatial: device tensors to route with instead of the workload's
    attributes (e.g. a KAN head's denormalised outputs). schedule_forcing:
    lay q' out in schedule order once here, as the data side of a batch
    would, so that the window calls permute nothing (False: the calls take
    q' in reach order and permute it themselves)."""

    def __init__(self, net, wl: Workload, qprime, gauges=None, window: Optional[int] = None,
                 params=None, n=None, q_spatial=None, schedule_forcing: bool = True):
        import torch
        from . import _native as NAT
        from .config import Params
        from .sparse import Workspace
        self.torch, self.NAT = torch, NAT
        self.net, self.wl, self.T, self.R = net, wl, int(qprime.shape[0]), wl.R
        self.qprime = qprime
        dev = qprime.device
        self.params = params if params is not None else Params()
        self.cp = self.params.to_c()
        self.window = window or (self.T - 1)
        self.ws = Workspace(net, self.window, self.T - 1, dense_grad=False,
                            schedule_forcing=schedule_forcing)
        # the q' rows the window calls read
        self.qwin = net.permute_forcing(qprime) if schedule_forcing else qprime
        f = lambda x: torch.as_tensor(np.ascontiguousarray(x, np.float32), device=dev)
        at = wl.attrs
        self.n = f(at["n"]) if n is None else n.detach().float().contiguous()
        self.qsp = f(at["q_spatial"]) if q_spatial is None else q_spatial.detach().float().contiguous()
        self.p = torch.full((self.R,), 21.0, dtype=torch.float32, device=dev)
        self.length, self.slope = f(at["length"]), f(at["slope"])
        sp = self.stream()
        NAT.check(NAT.lib.ddrs_mc_bind_statics(self.ws.handle, self.cp, self.n.data_ptr(),
                                               self.qsp.data_ptr(), self.p.data_ptr(),
                                               self.length.data_ptr(), self.slope.data_ptr(),
                                               None, sp), "ddrs_mc_bind_statics")
        self.q0 = torch.empty(self.R, dtype=torch.float32, device=dev)
        hot = net.workspace()
        NAT.check(NAT.lib.ddrs_hotstart(hot.handle, self.cp, qprime[0].data_ptr(),
                                        self.q0.data_ptr(), sp), "ddrs_hotstart")
        hot.check()
        self.gauges = np.asarray(wl.gauges if gauges is None else gauges, np.int32)
        self.ws.set_gauges(self.gauges)
        K = self.gauges.size
        self.Q = torch.empty((self.T, K), dtype=torch.float32, device=dev)
        self.gQ = torch.zeros((self.T, K), dtype=torch.float32, device=dev)
        self.gn = torch.empty(self.R, dtype=torch.float32, device=dev)
        self.gqsp = torch.empty(self.R, dtype=torch.float32, device=dev)

    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def forward(self, form=None, stream=None):
        NAT = self.NAT
        form = NAT.FORM_PERSISTENT if form is None else form
        NAT.check(NAT.lib.ddrs_mc_window_fwd_gauged(self.ws.handle, self.cp, self.T,
                                                    self.qwin.data_ptr(), self.q0.data_ptr(),
                                                    self.Q.data_ptr(), form,
                                                    stream or self.stream()),
                  "ddrs_mc_window_fwd_gauged")
        return self.Q

    def adjoint(self, form=None, stream=None):
        """Gradients of the loss whose dL/dQ at the gauges is self.gQ."""
        NAT = self.NAT
        form = NAT.FORM_PERSISTENT if form is None else form
        NAT.check(NAT.lib.ddrs_mc_window_adj_gauged(self.ws.handle, self.cp, self.T,
                                                    self.qwin.data_ptr(), self.q0.data_ptr(),
                                                    self.gQ.data_ptr(), self.gn.data_ptr(),
                                                    self.gqsp.data_ptr(), None, None, None,
                                                    form, stream or self.stream()),
                  "ddrs_mc_window_adj_gauged")
        return self.gn, self.gqsp

    def mse_grad(self, obs):
        """self.gQ = dL/dQ of L = mean over hours and gauges of (Q - obs)^2."""
        torch = self.torch
        scale = 2.0 / (self.T * self.gauges.size)
        torch.mul(torch.sub(self.Q, obs), scale, out=self.gQ)
        return self.gQ

    def step(self, obs, form=None, stream=None):
        """One bench step: forward, the MSE gradient at the gauges, adjoint."""
        self.forward(form, stream)
        self.mse_grad(obs)
        return self.adjoint(form, stream)

    def dense_forward(self, out=None, form=None):
        """The same forward writing discharge at every reach, [T, R]."""
        NAT = self.NAT
        form = NAT.FORM_PERSISTENT if form is None else form
        if out is None:
            out = self.torch.empty((self.T, self.R), dtype=self.torch.float32,
                                   device=self.qprime.device)
        NAT.check(NAT.lib.ddrs_mc_window_fwd(self.ws.handle, self.cp, self.T,
                                             self.qwin.data_ptr(), self.q0.data_ptr(),
                                             out.data_ptr(), form, self.stream()),
                  "ddrs_mc_window_fwd")
        return out

    def check(self):
        self.ws.check()

    def close(self):
        self.ws.close()

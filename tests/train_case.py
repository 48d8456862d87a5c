"""A nested-gauge training batch and the oracle chain of one micro-batch,
shared by the CPU test that pins the chain (finite differences) and the GPU
parity test that compares the device's training step with it.

The chain follows the reference's micro-batch (src/training/driver.rs:93-201,
src/training/forward.rs:301-387): KAN head (oracle/kan_oracle.py) ->
denormalize (src/routing/utils.rs:32-41) -> routing with the hotstart
(oracle/oracle.py) -> discharge at the gauges -> tau trim, daily pooling,
warmup, NaN filter and batch NSE (oracle/loss_oracle.py) -> the routing adjoint
with dL/dQ at the gauges -> the denormalize slope -> the KAN backward."""
import numpy as np

from oracle import kan_oracle as ko
from oracle import loss_oracle as lo
from oracle import oracle as O

F = 10
NAMES = ["n", "q_spatial"]
N_RANGE = (0.015, 0.25)
QSP_RANGE = (0.0, 1.0)
DENSE = ("input_weight", "input_bias", "output_weight", "output_bias")
BLOCK = ("coef", "scale_base", "scale_sp")


def oracle_params(head):
    """The head's weights (float64 numpy) in the KAN oracle's layout."""
    g = lambda n: getattr(head, n).detach().double().cpu().numpy()
    L = head.grid.shape[0]
    p = {n: g(n) for n in DENSE}
    p["blocks"] = [{n: g(n)[l] for n in ("grid", "coef", "scale_base", "scale_sp", "mask")}
                   for l in range(L)]
    return p


class Batch:
    """Nested gauges of log-uniform basin size on a synthetic tree, the union
    of their basins compressed to a batch, reach attributes ~ N(0, 1), storm
    forcing. observe() adds observations from a hidden head's routed
    discharge with 10% lognormal noise and one missing day."""

    def __init__(self, n_tree=4000, gauges=16, T=240, seed=0, tau=9, warmup=2, ddr_match=False):
        """ddr_match: the reference's DDR-faithful convention, both in the
        routing physics and in outflow_idx (a gauge's prediction sums its
        upstream reaches, data/collate.rs:310-326)."""
        from ddrs_amd import network as NW
        from ddrs_amd import train as TR
        rows, cols = NW.gen_tree(n_tree, 64, 500 + seed)
        g = TR.nested_gauges(n_tree, rows, cols, gauges, 5, 2000, seed=600 + seed)
        ur, uc = TR.upstream_union(n_tree, rows, cols, g)
        comp = NW.compress_ex(ur, uc, g, ddr_match=ddr_match)
        active, cr, cc, cg = comp.active, comp.rows, comp.cols, comp.gauge_compressed
        self.outflow = comp.scatter_map() if ddr_match else None
        R = active.size
        at = NW.gen_attributes(n_tree, 700 + seed)
        self.R, self.T, self.tau, self.warmup = R, T, tau, warmup
        self.rows, self.cols, self.gauges = cr, cc, cg
        self.length = at["length"][active]
        self.slope = at["slope"][active]
        rng = np.random.default_rng(800 + seed)
        self.x = rng.normal(size=(R, F)).astype(np.float32)
        self.qprime = NW.gen_forcing(R, T, 900 + seed, storms=True)
        self.P = O.make_params(ddr_match=ddr_match)
        self.pat = O.pattern_from_coo(R, cr, cc)
        self.obs = self.sigma = None

    def oracle_Qg(self, p, k, dtype=np.float32, nq=None):
        """Discharge at the gauges [T, G] for oracle head weights p, and what
        the backward needs. dtype: the routing's arithmetic (the KAN runs in
        float64 and is rounded to dtype). nq: (n, q_spatial) to route with
        instead of the oracle head's denormalised outputs."""
        out, cache = ko.kan_head_fwd(self.x, p, k)
        o = out.astype(dtype)
        if nq is None:
            n = O.denormalize(o[:, 0], *N_RANGE, False, dtype=dtype)
            qsp = O.denormalize(o[:, 1], *QSP_RANGE, False, dtype=dtype)
        else:
            n, qsp = (np.asarray(v, dtype) for v in nq)
        pp = np.full(self.R, 21.0, dtype)
        q0 = O.hotstart(self.P, self.pat, self.qprime[0], dtype=dtype)
        Q = O.route_fwd(self.P, self.pat, n, qsp, pp, self.length, self.slope, self.qprime,
                        q0=q0, dtype=dtype)
        return Q[:, self.readout()], (n, qsp, pp, q0, cache)

    def readout(self):
        """The reaches whose discharge forms the gauges' predictions."""
        return self.gauges if self.outflow is None else self.outflow[0]

    def scatter(self):
        """(flat, group) of scatter_add_by_group over the readout reaches."""
        if self.outflow is None:
            G = self.gauges.size
            return np.arange(G), np.arange(G)
        return self.outflow[1], self.outflow[2]

    def observe(self, p_truth, k):
        Qg, _ = self.oracle_Qg(p_truth, k)
        flat, grp = self.scatter()
        hourly = lo.scatter_add_by_group(Qg.T, flat, grp, self.gauges.size)
        daily = lo.tau_trim_and_downsample(hourly, self.tau)
        rng = np.random.default_rng(1)
        obs = (daily * np.exp(0.1 * rng.normal(size=daily.shape))).astype(np.float32)
        obs[3, self.warmup + 1] = np.nan
        self.obs = obs
        self.sigma = np.nanstd(obs, axis=1).astype(np.float32)

    def oracle_loss(self, p, k, eps=0.1, dtype=np.float32):
        Qg, _ = self.oracle_Qg(p, k, dtype)
        flat, grp = self.scatter()
        loss, nv, _ = lo.gauge_nse_loss(Qg.T, flat, grp, self.gauges.size, self.tau,
                                        self.warmup, self.obs, self.sigma, eps)
        return loss, nv

    def oracle_loss_and_grads(self, p, k, eps=0.1, dtype=np.float32, nq=None):
        """(loss, n_valid, {name: dL/dweight}) of the oracle chain (nq as in
        oracle_Qg)."""
        Qg, (n, qsp, pp, q0, cache) = self.oracle_Qg(p, k, dtype, nq)
        flat, grp = self.scatter()
        loss, nv, g_run = lo.gauge_nse_loss(Qg.T, flat, grp, self.gauges.size, self.tau,
                                            self.warmup, self.obs, self.sigma, eps)
        gQ = np.zeros((self.T, self.R), dtype)
        gQ[:, self.readout()] = g_run.T
        ga = O.route_adj(self.P, self.pat, n, qsp, pp, self.length, self.slope, self.qprime, gQ,
                         q0=q0, dtype=dtype)
        gout = np.stack([ga["n"].astype(np.float64) * (N_RANGE[1] - N_RANGE[0]),
                         ga["q_spatial"].astype(np.float64) * (QSP_RANGE[1] - QSP_RANGE[0])],
                        axis=1)
        g = ko.kan_head_bwd(gout, p, cache)
        grads = {n: g[n] for n in DENSE}
        for name in BLOCK:
            grads[name] = np.stack([b[name] for b in g["blocks"]])
        return loss, nv, grads

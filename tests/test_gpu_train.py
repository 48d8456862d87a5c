"""GPU parity of one training micro-batch (ddrs_amd.train.Trainer: KAN head ->
denormalize -> gauged routing window -> tau trim, daily pooling and batch NSE
-> routing adjoint -> KAN backward -> clipping -> AdaDelta) against the oracle
chain of tests/train_case.py (KAN oracle in float64, routing oracle in f32,
loss oracle in float64 after the f32 readout), which a CPU test pins to
float64 finite differences. Needs an MI355X.

Workload: a 4000-reach synthetic tree, 16 nested gauges of log-uniform basin
size, the union of their basins compressed to a batch (as the reference's
collate does), 240 hourly steps, observations from a hidden head's routed
discharge with 10% lognormal noise and one missing day (so one gauge is
dropped by the NaN filter).

The stages meet on the device's n and q_spatial: the oracle routes the
device head's denormalised outputs (and takes its KAN backward from its own
float64 forward). Fed its own head's outputs instead, which differ from the
device's by an ulp, the chain's weight gradients move by up to ~1e-3 max|g|:
perturbing the oracle's own n and q_spatial by one ulp moves its f32 weight
gradients by 7.5e-4 max|g| on this batch (reach states near the routing's
clamps, summed over reaches with 1-2% cancellation). That end-to-end
difference is printed, not asserted.

Bars:
* the loss within 1e-4 relative of the oracle chain, n_valid exact;
* each head gradient tensor within 1e-3 max|g| of the oracle chain (the
  routing gradient bar; the KAN and loss parts are far tighter);
* Trainer.step equals clip_grad_norm + AdaDelta applied on the CPU to the same
  device gradients (torch.optim.Adadelta, which the reference matches), within
  1e-6 relative;
* three more steps on the batch lower the loss.
"""
import sys

import numpy as np
import pytest
import torch

from tests import train_case as TC

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

import ddrs_amd as D  # noqa: E402
from ddrs_amd import kan as K  # noqa: E402
from ddrs_amd import train as TR  # noqa: E402

DEV = "cuda:0"


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def t(x):
    return torch.as_tensor(np.ascontiguousarray(x, np.float32), device=DEV)


def make_head(seed):
    return K.KanHeadConfig([f"a{i}" for i in range(TC.F)], TC.NAMES, seed).init(DEV)


def device_batch(b):
    return TR.GaugedBatch(b.adj_obj, t(b.x), t(b.qprime), b.gauges, t(b.obs), t(b.sigma),
                          outflow=b.outflow)


def make_batch(seed, ddr_match=False):
    from ddrs_amd.sparse import SparseAdjacency
    b = TC.Batch(seed=seed, ddr_match=ddr_match)
    b.observe(TC.oracle_params(make_head(99 + seed)), 3)
    b.adj_obj = SparseAdjacency.from_coo(b.R, b.rows, b.cols, None, b.length, b.slope)
    return b


@pytest.mark.parametrize("window,ddr_match", [(None, False), (50, False), (None, True)])
def test_micro_batch_matches_oracle_chain(window, ddr_match):
    """ddr_match: the reference's DDR-faithful convention in the physics and
    in outflow_idx (each gauge sums its upstream reaches, collate.rs:310-326),
    the batch built by compress_ex and the loss through scatter_add_by_group."""
    say("training micro-batch vs oracle chain, window", window, "ddr_match", ddr_match)
    b = make_batch(0, ddr_match)
    head = make_head(42)
    params = D.Params()
    params.ddr_match = ddr_match
    tr = TR.Trainer(head, params=params, tau=b.tau, warmup=b.warmup, window=window)
    batch = device_batch(b)
    loss, nv = tr.loss(batch)
    loss.backward()
    with torch.no_grad():
        out = head(batch.attributes)
        rng = tr.params.parameter_ranges
        nq = [(out[k] * (r[1] - r[0]) + r[0]).cpu().numpy()
              for k, r in (("n", rng.n), ("q_spatial", rng.q_spatial))]
    p = TC.oracle_params(head)
    wl, wnv, wg = b.oracle_loss_and_grads(p, head.k, nq=nq)
    _, _, own = b.oracle_loss_and_grads(p, head.k, dtype=np.float64)
    say(f"R={b.R} gauges={b.gauges.size}: loss {loss.item():.8g} oracle {wl:.8g}; "
        f"n_valid {int(nv)} oracle {wnv}")
    assert int(nv) == wnv == (b.gauges.size - 1) * ((b.T - 24) // 24 - b.warmup)
    assert abs(loss.item() - wl) <= 1e-4 * abs(wl)
    for name, want in wg.items():
        got = getattr(head, name).grad.double().cpu().numpy()
        scale = max(float(np.abs(want).max()), 1e-30)
        err = float(np.abs(got - want).max()) / scale
        e_own = float(np.abs(got - own[name]).max()) / float(np.abs(own[name]).max())
        print(f"  grad {name}: max err / max|g| {err:.3e} (vs the float64 chain through the "
              f"oracle head's own outputs {e_own:.3e})")
        assert err <= 1e-3, name


def test_step_is_clip_plus_adadelta_and_descends():
    say("training step = clip + AdaDelta; descent")
    b = make_batch(1)
    batch = device_batch(b)
    ha, hb = make_head(42), make_head(42)
    ta = TR.Trainer(ha, tau=b.tau, warmup=b.warmup)
    tb = TR.Trainer(hb, tau=b.tau, warmup=b.warmup, network=ta.network(batch))
    # reference: the device gradients, clipped and stepped on the CPU
    loss, _ = ta.loss(batch)
    loss.backward()
    cpu = [torch.nn.Parameter(p.detach().cpu().clone()) for p in ha.parameters()]
    for c, p in zip(cpu, ha.parameters()):
        c.grad = p.grad.detach().cpu().clone()
    norm = TR.clip_grad_norm(cpu, 1.0)
    torch.optim.Adadelta(cpu).step()
    out = tb.step(batch)
    assert out.grad_norm == pytest.approx(norm, rel=1e-6)
    assert out.loss == pytest.approx(loss.item(), rel=1e-7)
    for c, p in zip(cpu, hb.parameters()):
        np.testing.assert_allclose(p.detach().cpu().numpy(), c.detach().numpy(), rtol=1e-6,
                                   atol=1e-9)
    losses = [out.loss]
    for _ in range(3):
        losses.append(tb.step(batch).loss)
    say("losses", losses)
    assert losses[-1] < losses[0]


def test_adam_l1_step():
    """The reference's default objective and optimizer (L1, Adam at lr 1e-3):
    Trainer.step equals clipping + torch.optim.Adam on the CPU applied to the
    same device gradients, within 1e-6 relative; the L1 loss of the step
    equals the oracle chain's."""
    say("training step, L1 + Adam")
    b = make_batch(2)
    batch = TR.GaugedBatch(b.adj_obj, t(b.x), t(b.qprime), b.gauges, t(b.obs), None)
    ha, hb = make_head(42), make_head(42)
    ta = TR.Trainer(ha, tau=b.tau, warmup=b.warmup, loss="l1", optimizer="adam")
    tb = TR.Trainer(hb, tau=b.tau, warmup=b.warmup, loss="l1", optimizer="adam",
                    network=ta.network(batch))
    loss, nv = ta.loss(batch)
    loss.backward()
    cpu = [torch.nn.Parameter(p.detach().cpu().clone()) for p in ha.parameters()]
    for c, p in zip(cpu, ha.parameters()):
        c.grad = p.grad.detach().cpu().clone()
    norm = TR.clip_grad_norm(cpu, 1.0)
    torch.optim.Adam(cpu, lr=1e-3, betas=(0.9, 0.999), eps=1e-8).step()
    out = tb.step(batch)
    assert out.grad_norm == pytest.approx(norm, rel=1e-6)
    for c, p in zip(cpu, hb.parameters()):
        np.testing.assert_allclose(p.detach().cpu().numpy(), c.detach().numpy(), rtol=1e-6,
                                   atol=1e-9)
    with torch.no_grad():
        o = ha(batch.attributes)
        rng = ta.params.parameter_ranges
        nq = [(o[k] * (r[1] - r[0]) + r[0]).cpu().numpy()
              for k, r in (("n", rng.n), ("q_spatial", rng.q_spatial))]
    Qg, _ = b.oracle_Qg(TC.oracle_params(ha), ha.k, nq=nq)
    from oracle import loss_oracle as lo
    G = b.gauges.size
    wl, wnv, _ = lo.gauge_nse_loss(Qg.T, np.arange(G), np.arange(G), G, b.tau, b.warmup, b.obs,
                                   None, 0.1, kind="l1")
    assert int(nv) == wnv and abs(out.loss - wl) <= 1e-4 * abs(wl)


def test_gradient_accumulation_is_the_pooled_mean():
    """step_accumulated over two micro-batches of different networks: the
    stepped weights equal clipping + AdaDelta applied to
    (n1 g1 + n2 g2) / (n1 + n2) of the separate micro-batch gradients, within
    1e-6; the reported loss is the n_valid-weighted mean."""
    say("gradient accumulation over two micro-batches")
    b1, b2 = make_batch(3), make_batch(4)
    d1, d2 = device_batch(b1), device_batch(b2)
    ha, hb = make_head(42), make_head(42)
    ta = TR.Trainer(ha, tau=b1.tau, warmup=b1.warmup)
    grads, ns, losses = [], [], []
    for d in (d1, d2):
        for p in ha.parameters():
            p.grad = None
        loss, nv = ta.loss(d)
        loss.backward()
        grads.append([p.grad.detach().cpu().clone() for p in ha.parameters()])
        ns.append(int(nv))
        losses.append(loss.item())
    cpu = [torch.nn.Parameter(p.detach().cpu().clone()) for p in ha.parameters()]
    for i, c in enumerate(cpu):
        c.grad = (grads[0][i] * ns[0] + grads[1][i] * ns[1]) / (ns[0] + ns[1])
    TR.clip_grad_norm(cpu, 1.0)
    torch.optim.Adadelta(cpu).step()
    tb = TR.Trainer(hb, tau=b1.tau, warmup=b1.warmup)
    out = tb.step_accumulated([d1, d2])
    assert out.n_valid == sum(ns)
    assert out.loss == pytest.approx((losses[0] * ns[0] + losses[1] * ns[1]) / sum(ns),
                                     rel=1e-6)
    for c, p in zip(cpu, hb.parameters()):
        np.testing.assert_allclose(p.detach().cpu().numpy(), c.detach().numpy(), rtol=1e-6,
                                   atol=1e-9)


def test_learned_p_spatial_gets_its_gradient():
    """A head that also produces p_spatial routes with it (denormalised in log
    space, forward.rs:319-373): the chain's weight gradients match the oracle
    chain routed with the device head's three parameters, within 1e-3 max|g|."""
    say("learned p_spatial")
    b = make_batch(5)
    names = ["n", "q_spatial", "p_spatial"]
    head = K.KanHeadConfig([f"a{i}" for i in range(TC.F)], names, 42).init(DEV)
    tr = TR.Trainer(head, tau=b.tau, warmup=b.warmup)
    batch = device_batch(b)
    loss, nv = tr.loss(batch)
    loss.backward()
    import math
    from oracle import kan_oracle as ko
    from oracle import loss_oracle as lo
    from oracle import oracle as O
    with torch.no_grad():
        o = head(batch.attributes)
        rng = tr.params.parameter_ranges
        n = (o["n"] * (rng.n[1] - rng.n[0]) + rng.n[0]).cpu().numpy()
        q = (o["q_spatial"] * (rng.q_spatial[1] - rng.q_spatial[0]) + rng.q_spatial[0]).cpu().numpy()
        lo_, hi_ = math.log(rng.p_spatial[0] + 1e-6), math.log(rng.p_spatial[1])
        ps = torch.exp(o["p_spatial"] * (hi_ - lo_) + lo_).cpu().numpy()
    p = TC.oracle_params(head)
    out, cache = ko.kan_head_fwd(b.x, p, head.k)
    q0 = O.hotstart(b.P, b.pat, b.qprime[0])
    Q = O.route_fwd(b.P, b.pat, n, q, ps, b.length, b.slope, b.qprime, q0=q0)
    G = b.gauges.size
    wl, wnv, g_run = lo.gauge_nse_loss(Q[:, b.gauges].T, np.arange(G), np.arange(G), G, b.tau,
                                       b.warmup, b.obs, b.sigma, 0.1)
    gQ = np.zeros((b.T, b.R), np.float32)
    gQ[:, b.gauges] = g_run.T
    ga = O.route_adj(b.P, b.pat, n, q, ps, b.length, b.slope, b.qprime, gQ, q0=q0)
    gout = np.stack([ga["n"].astype(np.float64) * (rng.n[1] - rng.n[0]),
                     ga["q_spatial"].astype(np.float64),
                     ga["p_spatial"].astype(np.float64) * ps.astype(np.float64) * (hi_ - lo_)],
                    axis=1)
    g = ko.kan_head_bwd(gout, p, cache)
    assert int(nv) == wnv and abs(loss.item() - wl) <= 1e-4 * abs(wl)
    for name in ("input_weight", "output_weight", "output_bias"):
        got = getattr(head, name).grad.double().cpu().numpy()
        err = float(np.abs(got - g[name]).max() / np.abs(g[name]).max())
        print(f"  grad {name}: {err:.3e}")
        assert err <= 1e-3, name

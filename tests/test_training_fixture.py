"""The reference's training-step fixtures (tests/fixtures/training_step, dumped
from DDR for gauge 10336740: 19 reaches, a 2136-hour window, tau 3, warmup 5
days; copied to tests/golden/training_step_77006074.npz by
tests/golden/make_training_step_fixture.py) against the oracle and the host
optimizer, on the CPU:

* the KAN head at the seed-42 fixture weights on the batch's normalised
  attributes, denormalised (n and q_spatial linear, p_spatial in log space),
  gives the routed parameters of the fixture's forward;
* the L1 objective of its post-warmup prediction and observations
  (tests/training_step_layer_c.rs:378-420);
* the gradient norm before clipping (below 1: clipping leaves the gradients
  alone) and one Adam step from the fixture weights with the fixture
  gradients: parameters and both moments (tests/training_step_layer_d.rs).

The fixture's routed discharge cannot be reproduced here: its reach lengths
and slopes come from a CONUS store that the fixtures do not hold. Its daily
series is not compared either: DDR pools hours [13, 2128) into 88 days with
torch's adaptive average pooling (it matches that exactly), while the
reference's tau_trim_and_downsample, which the oracle restates and the
reference's own tests pin (src/training/loss.rs:609-727), pools hours
[tau, T - 24 + tau) with fractional area weights; the reference's test of this
fixture slices by hand with a convention that its current loss.rs no longer
uses.
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import kan_oracle as ko
from oracle import loss_oracle as lo
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
K_FIX = os.path.join(GOLD, "kan_head_seed42.npz")
T_FIX = os.path.join(GOLD, "training_step_77006074.npz")
DENSE = ("input_weight", "input_bias", "output_weight", "output_bias")
BLOCK = ("coef", "scale_base", "scale_sp")


def load(path):
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_kan_and_denormalize_give_the_routed_parameters():
    kf, tf = load(K_FIX), load(T_FIX)
    p = ko.params_from_npz(kf)
    out, _ = ko.kan_head_fwd(tf["mc_forward_norm_attrs"], p, 2)   # grid 50, order 2
    o = out.astype(np.float32)
    n = O.denormalize(o[:, 0], 0.015, 0.25, False)
    q = O.denormalize(o[:, 1], 0.0, 1.0, False)
    ps = O.denormalize(o[:, 2], 1.0, 200.0, True)
    for name, got in (("n", n), ("q_spatial", q), ("p_spatial", ps)):
        want = tf[f"mc_forward_{name}_param"]
        rel = float(np.abs(got - want).max() / np.abs(want).max())
        assert rel <= 1e-6, (name, rel)
    # the log-space branch matters: a linear p_spatial is far off
    lin = O.denormalize(o[:, 2], 1.0, 200.0, False)
    assert np.abs(lin - tf["mc_forward_p_spatial_param"]).max() > 1.0


def test_l1_objective():
    tf = load(T_FIX)
    p, o = tf["loss_and_grads_pred_post_warmup"], tf["loss_and_grads_obs_post_warmup"]
    assert p.shape == o.shape == (83,)
    assert abs(lo.l1_batch_loss(p, o) - tf["loss_and_grads_loss"].item()) <= 1e-6
    # the fused objective's oracle with kind l1 on the same series: hourly
    # values constant over each pooled day reproduce p exactly
    tau, warmup = 3, 5
    days = 88
    hourly = np.zeros((1, 24 * (days + 1)), np.float64)
    full_p = np.concatenate([np.zeros(warmup), p])
    for j in range(days):
        hourly[0, tau + 24 * j: tau + 24 * (j + 1)] = full_p[j]
    obs = np.concatenate([np.zeros(warmup), o])[None]
    loss, nv, _ = lo.gauge_nse_loss(hourly, [0], [0], 1, tau, warmup, obs, None, 0.1,
                                    kind="l1")
    assert nv == 83
    assert abs(loss - tf["loss_and_grads_loss"].item()) <= 1e-5


def _fixture_params_and_grads():
    kf, tf = load(K_FIX), load(T_FIX)
    names = list(DENSE) + [f"block_{l}_{b}" for l in range(2) for b in BLOCK]
    params = {k: torch.nn.Parameter(torch.tensor(kf[k], dtype=torch.float32)) for k in names}
    for k, prm in params.items():
        prm.grad = torch.tensor(tf[f"loss_and_grads_grad_{k}"], dtype=torch.float32)
    return names, params, tf


def test_grad_norm_and_clipping():
    from ddrs_amd import train as TR
    names, params, tf = _fixture_params_and_grads()
    before = {k: params[k].grad.clone() for k in names}
    norm = TR.clip_grad_norm([params[k] for k in names], 1.0)
    assert norm == pytest.approx(tf["loss_and_grads_post_clip_grad_norm"].item(), rel=1e-6)
    for k in names:
        assert torch.equal(params[k].grad, before[k])


def test_one_adam_step_matches_the_fixture():
    from ddrs_amd import train as TR
    names, params, tf = _fixture_params_and_grads()
    opt = TR.Adam([params[k] for k in names])
    opt.step(1e-3)
    for k in names:
        got = params[k].detach().numpy()
        want = tf[f"adam_step_{k}"]
        assert np.abs(got - want).max() <= 2e-7, k
        m, v = opt.moments(params[k])
        np.testing.assert_allclose(m.numpy(), tf[f"adam_step_moment1_{k}"], rtol=1e-6,
                                   atol=1e-12, err_msg=k)
        np.testing.assert_allclose(v.numpy(), tf[f"adam_step_moment2_{k}"], rtol=1e-6,
                                   atol=1e-18, err_msg=k)
    # the step moved every weight with a non-zero gradient by about lr
    d = np.abs(params["input_weight"].detach().numpy()
               - load(K_FIX)["input_weight"])
    assert math.isclose(float(np.median(d)), 1e-3, rel_tol=0.01)

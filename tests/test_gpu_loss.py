"""GPU parity of the gauge readout and training loss (ddrs_amd.loss through
the C ABI) against the reference's test values and the CPU oracle. Needs an
MI355X.

Bars:
* scatter_add_by_group and its backward: bit-exact against the float32
  oracle (both add in increasing k), in both series layouts;
* pooling: the reference's hand values (src/training/loss.rs:609-727) within
  their 1e-4 / 1e-5, and within 1e-6 relative of the oracle's float64
  application of the float32 weights;
* nse_batch_loss: the hand value of loss.rs:491-507 within 1e-6; the fused
  objective within 1e-5 relative of the float64 oracle, n_valid exact, the
  gradient within 1e-5 of its max-abs; deterministic bitwise.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

from oracle import loss_oracle as lo

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

from ddrs_amd import _native as NAT  # noqa: E402
from ddrs_amd import loss as L  # noqa: E402

DEV = "cuda:0"


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def t32(x):
    return torch.tensor(np.asarray(x, np.float32), device=DEV)


@pytest.mark.parametrize("time_major", [False, True])
def test_scatter_fixture(time_major):
    say("scatter fixture", time_major)
    runoff = np.array([[1, 10], [2, 20], [3, 30], [4, 40]], np.float32)
    plan = L.GaugePlan.from_outflow_idx(4, [[0, 1], [2], [3]], 2)
    x = t32(runoff.T.copy() if time_major else runoff)
    out = L.scatter_add_by_group(x, plan, time_major)
    assert out.cpu().numpy().ravel().tolist() == [3.0, 30.0, 3.0, 30.0, 4.0, 40.0]


@pytest.mark.parametrize("time_major", [False, True])
def test_scatter_random_bit_exact(time_major):
    say("scatter random", time_major)
    rng = np.random.default_rng(1)
    Nr, T, G, K = 500, 77, 40, 120
    flat = rng.integers(0, Nr, K)
    group = rng.integers(0, G - 1, K)          # gauge G-1 gets no reach
    runoff = rng.lognormal(0, 1, (Nr, T)).astype(np.float32)
    plan = L.GaugePlan(Nr, flat, group, G, T)
    x = t32(runoff.T.copy() if time_major else runoff).requires_grad_(True)
    out = L.scatter_add_by_group(x, plan, time_major)
    want = lo.scatter_add_by_group(runoff, flat, group, G)
    np.testing.assert_array_equal(out.detach().cpu().numpy(), want)
    gout = rng.normal(size=(G, T)).astype(np.float32)
    out.backward(t32(gout))
    g = x.grad.cpu().numpy()
    g = g.T if time_major else g
    np.testing.assert_array_equal(g, lo.scatter_add_by_group_bwd(gout, flat, group, Nr))


def test_pool_reference_values():
    say("pool reference values")
    out = L.tau_trim_and_downsample(t32(np.arange(72)[None]), 0).cpu().numpy()
    np.testing.assert_allclose(out[0], [11.5, 35.5], atol=1e-4)
    assert tuple(L.tau_trim_and_downsample(torch.zeros(1, 2160, device=DEV), 3).shape) == (1, 89)
    v = np.array([(x * 37) % 101 for x in range(240)], np.float32)
    new = L.tau_trim_and_downsample(t32(v[None]), 9).cpu().numpy()[0]
    for j in range(len(new) - 1):
        s = 33 + 24 * j
        assert abs(new[j + 1] - v[s:s + 24].mean()) < 1e-4


@pytest.mark.parametrize("t_hours,tau", [(2163, 3), (720, 0), (8760, 9), (1000, 23), (48, 5)])
def test_pool_matches_oracle_and_adjoint(t_hours, tau):
    say("pool", t_hours, tau)
    rng = np.random.default_rng(t_hours)
    x = rng.lognormal(0, 1, (7, t_hours)).astype(np.float32)
    xt = t32(x).requires_grad_(True)
    y = L.tau_trim_and_downsample(xt, tau)
    want = lo.tau_trim_and_downsample(x, tau)
    np.testing.assert_allclose(y.detach().cpu().numpy(), want, rtol=1e-6)
    g = rng.normal(size=want.shape).astype(np.float32)
    y.backward(t32(g))
    np.testing.assert_allclose(xt.grad.cpu().numpy(), lo.tau_trim_and_downsample_bwd(g, t_hours, tau),
                               rtol=1e-6, atol=1e-7)


def test_nse_hand_value_and_gradient_sign():
    say("nse hand value")
    plan = L.GaugePlan(1, [0], [0], 2, 24)
    p = t32([[2, 0, 2, 0], [5, 5, 5, 5]]).requires_grad_(True)
    o = t32([[1, 1, 1, 1], [3, 3, 3, 3]])
    v = L.nse_batch_loss(p, o, t32([1.0, 2.0]), 0.1, plan)
    want = (4 * 1 / 1.1 ** 2 + 4 * 4 / 2.1 ** 2) / 8
    assert abs(v.item() - want) < 1e-6
    v.backward()
    g = p.grad.cpu().numpy()
    assert (g[0, [0, 2]] > 0).all() and (g[0, [1, 3]] < 0).all()
    np.testing.assert_allclose(g, 2 * (p.detach().cpu().numpy() - o.cpu().numpy())
                               / (np.array([1.1, 2.1]) ** 2)[:, None] / 8, rtol=1e-6)


def _fused_case(seed, Nr, T, G, per_gauge, t_days_nan=()):
    rng = np.random.default_rng(seed)
    outflow = [list(rng.choice(Nr, per_gauge, replace=False)) for _ in range(G)]
    outflow[0].append(outflow[1][0])           # a reach shared by two gauges
    flat = [int(s) for o in outflow for s in o]
    group = [g for g, o in enumerate(outflow) for _ in o]
    runoff = rng.lognormal(0, 0.7, (Nr, T)).astype(np.float32)
    t_days = (T - 24) // 24
    obs = rng.lognormal(0, 0.7, (G, t_days)).astype(np.float32)
    for g, d in t_days_nan:
        obs[g, d] = np.nan
    sigma = rng.uniform(0.3, 2.0, G).astype(np.float32)
    return flat, group, runoff, obs, sigma


@pytest.mark.parametrize("time_major", [True, False])
def test_fused_objective_matches_oracle(time_major):
    say("fused objective", time_major)
    Nr, T, G = 3000, 24 * 40 + 7, 25
    flat, group, runoff, obs, sigma = _fused_case(5, Nr, T, G, 3,
                                                  [(3, 10), (7, 0), (11, 39 - 1), (2, 2)])
    tau, warmup = 9, 3
    plan = L.GaugePlan(Nr, flat, group, G, T)
    x = t32(runoff.T.copy() if time_major else runoff).requires_grad_(True)
    loss, n_valid, hourly = L.gauge_nse_loss(x, plan, tau, warmup, t32(obs), t32(sigma), 0.1,
                                             time_major)
    loss.backward()
    wl, wn, wg = lo.gauge_nse_loss(runoff, flat, group, G, tau, warmup, obs, sigma, 0.1)
    say(f"loss {loss.item():.8g} oracle {wl:.8g}, n_valid {int(n_valid)} oracle {wn}")
    assert int(n_valid) == wn == (G - 2) * ((T - 24) // 24 - warmup)  # gauges 3 and 11 dropped
    assert abs(loss.item() - wl) <= 1e-5 * abs(wl)
    np.testing.assert_array_equal(hourly.cpu().numpy(), lo.scatter_add_by_group(runoff, flat, group, G))
    g = x.grad.cpu().numpy()
    g = g.T if time_major else g
    assert np.abs(g - wg).max() <= 1e-5 * np.abs(wg).max()
    untouched = np.setdiff1d(np.arange(Nr), flat)
    assert not g[untouched].any()
    # deterministic
    x2 = t32(runoff.T.copy() if time_major else runoff).requires_grad_(True)
    l2, _, _ = L.gauge_nse_loss(x2, plan, tau, warmup, t32(obs), t32(sigma), 0.1, time_major)
    l2.backward()
    assert l2.item() == loss.item()
    assert torch.equal(x2.grad, x.grad)


def test_fused_all_gauges_dropped_and_errors():
    say("fused all dropped, errors")
    Nr, T, G = 50, 24 * 5, 4
    flat, group, runoff, obs, sigma = _fused_case(6, Nr, T, G, 2)
    obs[:, -1] = np.nan
    plan = L.GaugePlan(Nr, flat, group, G, T)
    x = t32(runoff.T.copy()).requires_grad_(True)
    loss, n_valid, _ = L.gauge_nse_loss(x, plan, 0, 1, t32(obs), t32(sigma))
    loss.backward()
    assert loss.item() == 0.0 and int(n_valid) == 0 and not x.grad.any()
    with pytest.raises(NAT.DdrsError):
        L.gauge_nse_loss(x, plan, 24, 0, t32(obs), t32(sigma))        # tau outside [0, 24)
    with pytest.raises(NAT.DdrsError):
        L.gauge_nse_loss(x, plan, 0, 4, t32(obs), t32(sigma))         # warmup >= 4 days
    big = torch.zeros(T + 24, Nr, device=DEV)
    with pytest.raises(NAT.DdrsError):
        L.gauge_nse_loss(big, plan, 0, 0, t32(np.zeros((G, 5))), t32(sigma))  # > max_t_hours
    with pytest.raises(NAT.DdrsError):
        L.GaugePlan(Nr, [0, Nr], [0, 1], G, T)                        # reach out of range
    with pytest.raises(NAT.DdrsError):
        L.GaugePlan(Nr, [0, 1], [0, G], G, T)                         # gauge out of range


def test_workload_readout_c2_shape():
    """Time-major discharge of the 350,000-reach, 720-step workload with 64
    gauges (one outflow reach each, as the bench's loss reads them)."""
    Nr, T, G = 350_000, 720, 64
    say(f"fused objective at {T} x {Nr}, {G} gauges")
    rng = np.random.default_rng(8)
    reaches = np.sort(np.concatenate([[Nr - 1], rng.choice(Nr - 1, G - 1, replace=False)]))
    plan = L.GaugePlan(Nr, reaches, np.arange(G), G, T)
    Q = torch.rand(T, Nr, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3)) + 0.1
    Qh = Q[:, torch.as_tensor(reaches, device=DEV)].T.cpu().numpy()   # [G, T]
    obs = rng.lognormal(0, 0.3, (G, (T - 24) // 24)).astype(np.float32)
    sigma = np.full(G, 0.5, np.float32)
    x = Q.clone().requires_grad_(True)
    loss, n_valid, hourly = L.gauge_nse_loss(x, plan, 0, 2, t32(obs), t32(sigma), 0.1)
    loss.backward()
    np.testing.assert_array_equal(hourly.cpu().numpy(), Qh)
    wl, wn, wg = lo.gauge_nse_loss(Qh, list(range(G)), list(range(G)), G, 0, 2, obs, sigma, 0.1)
    assert int(n_valid) == wn and abs(loss.item() - wl) <= 1e-5 * wl
    g = x.grad[:, torch.as_tensor(reaches, device=DEV)].T.cpu().numpy()
    assert np.abs(g - wg).max() <= 1e-5 * np.abs(wg).max()
    assert float(x.grad.abs().sum()) == pytest.approx(float(np.abs(g).sum()), rel=1e-6)


@pytest.mark.parametrize("time_major", [True, False])
def test_fused_l1_matches_oracle(time_major):
    """The L1 kind of the fused objective (loss.rs:186, the reference's default
    loss): loss within 1e-5 relative of the float64 oracle, n_valid exact, the
    gradient sign(p - o) / n_valid through the pooling and readout within 1e-5
    of its max-abs; sigma may be absent."""
    say("fused L1", time_major)
    Nr, T, G = 2000, 24 * 30 + 5, 17
    flat, group, runoff, obs, sigma = _fused_case(9, Nr, T, G, 2, [(4, 6)])
    tau, warmup = 3, 2
    plan = L.GaugePlan(Nr, flat, group, G, T)
    x = t32(runoff.T.copy() if time_major else runoff).requires_grad_(True)
    loss, n_valid, _ = L.gauge_batch_loss(x, plan, tau, warmup, t32(obs), None, kind="l1",
                                          time_major=time_major)
    loss.backward()
    wl, wn, wg = lo.gauge_nse_loss(runoff, flat, group, G, tau, warmup, obs, None, 0.1,
                                   kind="l1")
    say(f"loss {loss.item():.8g} oracle {wl:.8g}")
    assert int(n_valid) == wn == (G - 1) * ((T - 24) // 24 - warmup)
    assert abs(loss.item() - wl) <= 1e-5 * abs(wl)
    g = x.grad.cpu().numpy()
    g = g.T if time_major else g
    assert np.abs(g - wg).max() <= 1e-5 * np.abs(wg).max()
    with pytest.raises(ValueError):
        L.gauge_batch_loss(x, plan, tau, warmup, t32(obs), None, kind="nse-batch")


@pytest.mark.parametrize("kind,weights", [
    ("nnse-kge", {}),
    ("nnse-kge", dict(nnse_weight=0.5, kge_weight=2.0)),
    ("kge", {}),
    ("kge", dict(nnse_weight=0.0, alpha_weight=2.0, kge_clamp=0.3)),
])
@pytest.mark.parametrize("time_major", [True, False])
def test_fused_per_gauge_kinds_match_oracle(kind, weights, time_major):
    """The per-gauge objectives nnse-kge and kge (loss.rs:257-373) through the
    fused call: loss within 1e-5 relative of the float64 oracle, n_valid the
    kept gauges, the gradient within 1e-5 of its max-abs."""
    say("fused", kind, weights, time_major)
    Nr, T, G = 1500, 24 * 20 + 7, 9
    flat, group, runoff, obs, sigma = _fused_case(12, Nr, T, G, 2, [(5, 8)])
    tau, warmup = 9, 3
    plan = L.GaugePlan(Nr, flat, group, G, T)
    x = t32(runoff.T.copy() if time_major else runoff).requires_grad_(True)
    loss, n_valid, _ = L.gauge_batch_loss(x, plan, tau, warmup, t32(obs), None, 0.1, kind=kind,
                                          time_major=time_major, **weights)
    loss.backward()
    ow = {("clamp" if k == "kge_clamp" else k): v for k, v in weights.items()}
    wl, wn, wg = lo.gauge_nse_loss(runoff, flat, group, G, tau, warmup, obs, None, 0.1,
                                   kind=kind, **ow)
    say(f"loss {loss.item():.8g} oracle {wl:.8g}")
    assert int(n_valid) == wn == G - 1
    assert abs(loss.item() - wl) <= 1e-5 * abs(wl)
    g = x.grad.cpu().numpy()
    g = g.T if time_major else g
    assert np.abs(g - wg).max() <= 1e-5 * np.abs(wg).max()


def test_per_gauge_kinds_reference_hand_value():
    """nnse_kge_loss's hand value (loss.rs:391-403: 0.7 at eps 0) through the
    fused call: hourly values constant over each pooled day give the daily
    series [1.5, 2.5, 1.5, 2.5] against observations [1, 3, 1, 3]."""
    say("per-gauge kinds hand value")
    days, tau = 4, 0
    hourly = np.zeros((1, 24 * (days + 1)), np.float32)
    for j, v in enumerate([1.5, 2.5, 1.5, 2.5]):
        hourly[0, 24 * j: 24 * (j + 1)] = v
    plan = L.GaugePlan(1, [0], [0], 1, hourly.shape[1])
    obs = t32([[1.0, 3.0, 1.0, 3.0]])
    x = t32(hourly)
    loss, nv, _ = L.gauge_batch_loss(x, plan, tau, 0, obs, None, 0.0, kind="nnse-kge",
                                     time_major=False)
    assert abs(loss.item() - 0.7) < 1e-6 and int(nv) == 1
    loss, _, _ = L.gauge_batch_loss(x, plan, tau, 0, obs, None, 0.0, kind="kge",
                                    time_major=False, nnse_weight=0.0, kge_clamp=1e9)
    assert abs(loss.item() - 0.25) < 1e-6

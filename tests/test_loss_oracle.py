"""The loss-side oracle (oracle/loss_oracle.py) pinned on the CPU against the
reference's own test values (src/training/forward.rs:1030-1058,
src/training/loss.rs:491-727), torch's adaptive average pooling where the
trimmed window is a whole number of days, and central differences."""
import numpy as np
import pytest
import torch

from oracle import loss_oracle as lo


def test_scatter_add_fixture():
    # forward.rs:1030-1058: 4 segments x 2 steps, outflow_idx [[0, 1], [2], [3]]
    runoff = np.array([[1, 10], [2, 20], [3, 30], [4, 40]], np.float32)
    out = lo.scatter_add_by_group(runoff, [0, 1, 2, 3], [0, 0, 1, 2], 3)
    assert out.ravel().tolist() == [3.0, 30.0, 3.0, 30.0, 4.0, 40.0]


def test_scatter_add_duplicates_and_unordered_groups():
    rng = np.random.default_rng(0)
    runoff = rng.normal(size=(7, 5)).astype(np.float32)
    flat = [3, 1, 3, 6, 0, 1]
    group = [2, 0, 1, 2, 0, 1]
    out = lo.scatter_add_by_group(runoff, flat, group, 4)
    want = np.zeros((4, 5))
    for f, g in zip(flat, group):
        want[g] += runoff[f]
    np.testing.assert_allclose(out, want, rtol=1e-6)
    assert not out[3].any()
    gout = rng.normal(size=(4, 5))
    g = lo.scatter_add_by_group_bwd(gout, flat, group, 7)
    # adjoint identity <gout, A x> = <A^T gout, x>
    np.testing.assert_allclose((gout * out).sum(), (g * runoff).sum(), rtol=1e-5)


def test_nse_batch_loss_hand_value():
    # loss.rs:491-507
    p = np.array([[2, 0, 2, 0], [5, 5, 5, 5]], np.float32)
    o = np.array([[1, 1, 1, 1], [3, 3, 3, 3]], np.float32)
    v = lo.nse_batch_loss(p, o, [1.0, 2.0], 0.1)
    want = (4 * 1 / 1.1 ** 2 + 4 * 4 / 2.1 ** 2) / 8
    assert abs(v - want) < 1e-5
    assert abs(v - 0.866738) < 1e-5


def test_area_pool_rows_sum_to_one_and_non_divisible_row():
    # loss.rs:609-665
    W = lo.area_pool_weights(2139, 89)
    np.testing.assert_allclose(W.sum(axis=1), 1.0, atol=1e-5)
    s = np.float32(2139) / np.float32(89)
    np.testing.assert_allclose(W[0, :24], 1.0 / s, atol=1e-6)
    assert abs(W[0, 24] - (s - 24) / s) < 1e-4
    assert not W[0, 25:].any()
    # at most ceil(L/M) + 1 non-zeros per row
    assert ((W > 0).sum(axis=1) <= int(np.ceil(2139 / 89)) + 1).all()


def test_block_mean_cases():
    # loss.rs:619-637: 1..48 into 2 bins; loss.rs:676-695: tau 0 over 72 hours
    out = np.arange(1, 49, dtype=np.float32)[None] @ lo.area_pool_weights(48, 2).T
    np.testing.assert_allclose(out[0], [12.5, 36.5], atol=1e-5)
    d = lo.tau_trim_and_downsample(np.arange(72, dtype=np.float32)[None], 0)
    np.testing.assert_allclose(d[0], [11.5, 35.5], atol=1e-4)
    assert lo.tau_trim_and_downsample(np.zeros((1, 2160), np.float32), 3).shape == (1, 89)


def test_legacy_tau_equivalence():
    # loss.rs:697-727: new tau 9 day j+1 == hours [33 + 24 j, 57 + 24 j) mean
    v = np.array([(x * 37) % 101 for x in range(240)], np.float32)
    new = lo.tau_trim_and_downsample(v[None], 9)[0]
    for j in range(len(new) - 1):
        s = 33 + 24 * j
        assert abs(new[j + 1] - v[s:s + 24].mean()) < 1e-4


@pytest.mark.parametrize("t_hours,tau", [(2160, 3), (720, 0), (8760, 9), (96, 23)])
def test_pool_matches_torch_adaptive_pool_on_whole_days(t_hours, tau):
    rng = np.random.default_rng(t_hours + tau)
    x = rng.lognormal(0, 1, (5, t_hours)).astype(np.float32)
    start, end, L, M = lo.trim_window(t_hours, tau)
    assert L % 24 == 0
    ref = torch.nn.functional.adaptive_avg_pool1d(
        torch.tensor(x[:, start:end], dtype=torch.float64)[:, None], M)[:, 0].numpy()
    # the weights are float32 (1/24 rounds at 3e-9 relative)
    np.testing.assert_allclose(lo.tau_trim_and_downsample(x, tau), ref, rtol=1e-7)


def test_pool_backward_is_the_adjoint():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(3, 100)).astype(np.float32)   # 76 trimmed hours, 3 bins
    y = lo.tau_trim_and_downsample(x, 5)
    g = rng.normal(size=y.shape)
    gx = lo.tau_trim_and_downsample_bwd(g, 100, 5)
    np.testing.assert_allclose((g * y).sum(), (gx * x).sum(), rtol=1e-10)
    assert not gx[:, :5].any() and not gx[:, 100 - 19:].any()


def test_gauge_loss_matches_central_differences_and_filter():
    rng = np.random.default_rng(9)
    N, T, G = 6, 24 * 6 + 5, 3
    runoff = rng.lognormal(0, 0.5, (N, T))
    flat, group = [0, 2, 2, 5, 4], [0, 0, 1, 1, 2]
    t_days = (T - 24) // 24
    obs = rng.lognormal(0, 0.5, (G, t_days))
    obs[2, 4] = np.nan            # gauge 2 dropped (NaN after warmup)
    obs[0, 0] = np.nan            # inside the warmup: gauge 0 kept
    sigma = np.array([0.5, 1.5, 2.0])
    args = (flat, group, G, 7, 1, obs, sigma, 0.1)
    loss, n_valid, g = lo.gauge_nse_loss(runoff, *args)
    assert n_valid == 2 * (t_days - 1)
    assert not g[4].any()         # only the dropped gauge reads reach 4
    eps = 1e-6
    for idx in [(0, 30), (2, 7), (2, 100), (5, 50), (1, 40)]:
        rp, rm = runoff.copy(), runoff.copy()
        rp[idx] += eps
        rm[idx] -= eps
        fd = (lo.gauge_nse_loss(rp, *args)[0] - lo.gauge_nse_loss(rm, *args)[0]) / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-6 + 1e-4 * abs(g[idx]), idx
    # every gauge dropped: loss 0 with n_valid 0
    obs_nan = obs.copy()
    obs_nan[:, -1] = np.nan
    assert lo.gauge_nse_loss(runoff, flat, group, G, 7, 1, obs_nan, sigma, 0.1)[:2] == (0.0, 0)


def test_kge_objectives_reference_values():
    """nnse_kge_loss and kge_component_loss at the reference's hand values and
    properties (src/training/loss.rs:391-490)."""
    p = np.array([[1.5, 2.5, 1.5, 2.5]])
    o = np.array([[1.0, 3.0, 1.0, 3.0]])
    # 1 - KGE = 0.5 (alpha 0.5), 1 - NNSE = 0.2
    assert abs(lo.kge_batch_loss(p, o, "nnse-kge", 0.0)[0] - 0.7) < 1e-12
    # perfect prediction: 0; attenuated: larger
    assert lo.kge_batch_loss(o, o, "nnse-kge", 0.0, nnse_weight=0.0)[0] < 1e-12
    assert lo.kge_batch_loss(p, o, "nnse-kge", 0.0, nnse_weight=0.0)[0] > 0.1
    # the components: (0.5 - 1)^2 = 0.25, doubled by alpha_weight 2
    l1 = lo.kge_batch_loss(p, o, "kge", 0.0, nnse_weight=0.0, clamp=1e9)[0]
    l2 = lo.kge_batch_loss(p, o, "kge", 0.0, nnse_weight=0.0, alpha_weight=2.0, clamp=1e9)[0]
    assert abs(l1 - 0.25) < 1e-12 and abs(l2 - 2 * l1) < 1e-12
    # a collapsed-variance gauge blows up without the clamp, not with it
    p2 = np.array([[1.5, 2.5, 1.5, 2.5], [10.0, 90.0, 10.0, 90.0]])
    o2 = np.array([[1.0, 3.0, 1.0, 3.0], [1.0, 1.0, 1.0, 1.0001]])
    assert lo.kge_batch_loss(p2, o2, "kge", 1e-6, nnse_weight=0.0, alpha_weight=2.0,
                             clamp=1e9)[0] > 100.0
    assert lo.kge_batch_loss(p2, o2, "kge", 1e-6, nnse_weight=0.0, alpha_weight=2.0,
                             clamp=5.0)[0] <= 5.0 + 1e-9
    # gradients push the attenuated peaks (1, 3) up and the troughs (0, 2) down
    for kind in ("nnse-kge", "kge"):
        _, g = lo.kge_batch_loss(p, o, kind, 0.1, nnse_weight=0.0)
        assert g[0, 1] < 0 and g[0, 3] < 0 and g[0, 0] > 0 and g[0, 2] > 0


@pytest.mark.parametrize("kind,weights", [
    ("nnse-kge", dict(nnse_weight=0.7, kge_weight=1.3)),
    ("kge", dict(nnse_weight=0.5, r_weight=1.2, alpha_weight=2.0, beta_weight=0.8, clamp=10.0)),
    ("kge", dict(nnse_weight=0.0, clamp=0.05)),
])
def test_kge_gradients_against_finite_differences(kind, weights):
    rng = np.random.default_rng(3)
    P = rng.lognormal(0, 0.5, (3, 9))
    O = rng.lognormal(0, 0.5, (3, 9))
    _, g = lo.kge_batch_loss(P, O, kind, 0.1, **weights)
    h = 1e-6
    for i in range(3):
        for j in range(9):
            Pp, Pm = P.copy(), P.copy()
            Pp[i, j] += h
            Pm[i, j] -= h
            fd = (lo.kge_batch_loss(Pp, O, kind, 0.1, **weights)[0]
                  - lo.kge_batch_loss(Pm, O, kind, 0.1, **weights)[0]) / (2 * h)
            assert abs(fd - g[i, j]) <= 1e-7 * max(1.0, np.abs(g).max()), (i, j, fd, g[i, j])


def test_gauge_objective_per_gauge_kinds_count_gauges():
    rng = np.random.default_rng(4)
    runoff = rng.lognormal(0, 0.5, (6, 24 * 8)).astype(np.float32)
    obs = rng.lognormal(0, 0.5, (3, 7)).astype(np.float32)
    obs[1, 4] = np.nan
    flat, group = [0, 1, 2, 3, 4, 5], [0, 0, 1, 1, 2, 2]
    for kind in ("nnse-kge", "kge"):
        loss, nv, g = lo.gauge_nse_loss(runoff, flat, group, 3, 5, 2, obs, None, 0.1, kind=kind)
        assert nv == 2 and np.isfinite(loss) and not g[2:4].any() and g[[0, 1, 4, 5]].any()

"""The KAN-head oracle (oracle/kan_oracle.py) pinned on the CPU.

- Against the reference's fixture tests/fixtures/kan_head_init_seed42.npz
  (numeric copy in tests/golden/kan_head_seed42.npz): forward outputs to 1e-6
  and the gradients of sum(all outputs) to 1e-5, the tolerances of the
  reference's tests/kan_head_fixture_forward.rs:69 and
  tests/kan_head_fixture_backward.rs:77.
- Against what the mathematics fixes: the B-spline basis is a partition of
  unity on the inner knot range, order 1 gives the hat functions, and the
  reverse mode agrees with central differences (this also covers dL/d inputs,
  which the fixture does not hold).
"""
import os

import numpy as np
import pytest

from oracle import kan_oracle as ko

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "kan_head_seed42.npz")
NAMES = ("n", "q_spatial", "p_spatial")


def _fixture():
    with np.load(FIXTURE, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_matches_fixture(dtype):
    d = _fixture()
    p = ko.params_from_npz(d, 2)
    out, _ = ko.kan_head_fwd(d["inputs"], p, k=2, dtype=dtype)
    for j, name in enumerate(NAMES):
        np.testing.assert_allclose(out[:, j], d[f"expected_{name}"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_backward_matches_fixture(dtype):
    d = _fixture()
    p = ko.params_from_npz(d, 2)
    out, cache = ko.kan_head_fwd(d["inputs"], p, k=2, dtype=dtype)
    g = ko.kan_head_bwd(np.ones_like(out), p, cache, dtype=dtype)
    for name in ("input_weight", "input_bias", "output_weight", "output_bias"):
        np.testing.assert_allclose(g[name], d[f"grad_{name}"], rtol=0, atol=1e-5, err_msg=name)
    for l in range(2):
        for name in ("coef", "scale_base", "scale_sp"):
            np.testing.assert_allclose(g["blocks"][l][name], d[f"grad_block_{l}_{name}"], rtol=0,
                                       atol=1e-5, err_msg=f"block {l} {name}")


def _uniform_grid(num, k, n_in):
    h = 2.0 / num
    return np.tile(np.arange(-k, num + k + 1) * h - 1.0, (n_in, 1))


@pytest.mark.parametrize("k,num", [(1, 4), (2, 50), (3, 5), (5, 7)])
def test_basis_partition_of_unity(k, num):
    grid = _uniform_grid(num, k, 1)
    rng = np.random.default_rng(k)
    # on [g_k, g_{G-1-k}) every x has k+1 supporting functions summing to 1
    lo, hi = grid[0, k], grid[0, -1 - k]
    x = np.concatenate([rng.uniform(lo, hi, 500), grid[0, k:-1 - k]])[:, None]
    B, dB = ko.bspline_basis(x, grid, k)
    np.testing.assert_allclose(B.sum(axis=2), 1.0, atol=1e-12)
    np.testing.assert_allclose(dB.sum(axis=2), 0.0, atol=1e-9)
    assert (B >= -1e-15).all()
    assert ((B > 0).sum(axis=2) <= k + 1).all()
    # outside the knot range every function is zero
    out = np.array([[grid[0, 0] - 0.1], [grid[0, -1]], [grid[0, -1] + 3.0]])
    Bo, dBo = ko.bspline_basis(out, grid, k)
    assert not Bo.any() and not dBo.any()


def test_basis_order_one_is_hat_functions():
    grid = _uniform_grid(4, 1, 1)          # knots -1.5, -1, ..., 1.5
    x = np.linspace(-1.6, 1.6, 97)[:, None]
    B, _ = ko.bspline_basis(x, grid, 1)
    g = grid[0]
    for c in range(B.shape[2]):
        # hat on [g_c, g_{c+2}] peaking at g_{c+1}
        hat = np.clip(1.0 - np.abs(x[:, 0] - g[c + 1]) / (g[c + 1] - g[c]), 0.0, None)
        hat[(x[:, 0] < g[c]) | (x[:, 0] >= g[c + 2])] = 0.0
        np.testing.assert_allclose(B[:, 0, c], hat, atol=1e-12, err_msg=f"c={c}")


def _random_head(rng, F, H, P, L, num, k):
    G = num + 2 * k + 1
    p = dict(input_weight=rng.normal(0, 0.6, (H, F)), input_bias=rng.normal(0, 0.1, H),
             output_weight=rng.normal(0, 0.5, (P, H)), output_bias=rng.normal(0, 0.1, P),
             blocks=[])
    for _ in range(L):
        p["blocks"].append(dict(grid=_uniform_grid(num, k, H),
                                coef=rng.normal(0, 0.3, (H, H, G - 1 - k)),
                                scale_base=rng.uniform(-0.3, 0.3, (H, H)),
                                scale_sp=rng.uniform(0.1, 0.4, (H, H)),
                                mask=(rng.uniform(size=(H, H)) > 0.2).astype(float)))
    return p


@pytest.mark.parametrize("k,num", [(2, 6), (3, 5)])
def test_backward_matches_central_differences(k, num):
    rng = np.random.default_rng(7)
    F, H, P, L, N = 4, 5, 3, 2, 6
    p = _random_head(rng, F, H, P, L, num, k)
    x = rng.uniform(-1.5, 1.5, (N, F))
    w = rng.normal(size=(N, P))                  # L = sum(w * out)

    def loss(pp, xx):
        return float((ko.kan_head_fwd(xx, pp, k)[0] * w).sum())

    out, cache = ko.kan_head_fwd(x, p, k)
    g = ko.kan_head_bwd(w, p, cache)
    eps = 1e-6

    def fd(get, set_, shape):
        res = np.zeros(shape)
        for idx in np.ndindex(*shape):
            v = get()[idx]
            set_(idx, v + eps)
            lp = loss(p, x)
            set_(idx, v - eps)
            lm = loss(p, x)
            set_(idx, v)
            res[idx] = (lp - lm) / (2 * eps)
        return res

    for name in ("input_weight", "input_bias", "output_weight", "output_bias"):
        arr = p[name]
        num_g = fd(lambda: arr, lambda i, v: arr.__setitem__(i, v), arr.shape)
        np.testing.assert_allclose(g[name], num_g, rtol=1e-5, atol=1e-7, err_msg=name)
    for l in range(L):
        for name in ("coef", "scale_base", "scale_sp"):
            arr = p["blocks"][l][name]
            num_g = fd(lambda: arr, lambda i, v: arr.__setitem__(i, v), arr.shape)
            np.testing.assert_allclose(g["blocks"][l][name], num_g, rtol=1e-5, atol=1e-7,
                                       err_msg=f"block {l} {name}")
    num_gx = np.zeros_like(x)
    for idx in np.ndindex(*x.shape):
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        num_gx[idx] = (loss(p, xp) - loss(p, xm)) / (2 * eps)
    np.testing.assert_allclose(g["inputs"], num_gx, rtol=1e-5, atol=1e-7)

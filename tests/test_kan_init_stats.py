"""KanHeadConfig.init against the reference's DDR init statistics
(tests/fixtures/kan_init_stats_ddr.csv, copied to tests/golden by
tests/golden/make_kan_init_stats_fixture.py: a DDR head with 10 attributes,
3 outputs, hidden 21, 2 blocks, grid 50, order 2). The draws come from torch's
generator, not DDR's, so values differ and only the distributions can agree.
On the CPU:

* the reference's own check (tests/kan_head_init_parity.rs:56-82,111-222) at
  seed 42: relative error of mean and std within 6 %, an absolute bar of 0.05
  for values below 0.1, exact zeros for zeros;
* a stricter one where that bar says nothing (every std below 0.1): DDR's
  sample mean and std of each random tensor lie within 4 standard deviations
  of the spread of ours over 24 seeds, and our pooled std matches the
  distribution's closed form (Kaiming sqrt(2 / F), Xavier 0.1 sqrt(2 / (H + P)),
  U(-1, 1) / sqrt(H) -> 1 / sqrt(3 H));
* the deterministic tensors exactly: the extended grid (std, min, max), the
  constant scale_sp = 1 / sqrt(H), mask 1, zero biases.
The coefficient statistics pin the fit's convention: 51 noise values fitted
with 52 order-2 B-splines is underdetermined, and only the minimum-norm
least-squares solution reproduces DDR's std (0.0101 for noise of std 0.0017).
"""
import csv
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ddrs_amd.kan import KanHeadConfig  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kan_init_stats_ddr.csv")
F, H, P, GRID, K = 10, 21, 3, 50, 2
SEEDS = range(24)


def ddr_stats():
    with open(GOLD, newline="") as f:
        return {r["name"]: (float(r["mean"]), float(r["std"]), float(r["min"]), float(r["max"]))
                for r in csv.DictReader(f)}


def head_tensors(seed):
    cfg = KanHeadConfig([f"attr{i}" for i in range(F)], ["n", "q_spatial", "p_spatial"], seed,
                        hidden_size=H, num_hidden_layers=2, grid=GRID, k=K)
    sd = {k: v.detach().double().cpu().numpy() for k, v in cfg.init("cpu").state_dict().items()}
    out = {"input.weight": sd["input_weight"], "input.bias": sd["input_bias"],
           "output.weight": sd["output_weight"], "output.bias": sd["output_bias"]}
    for b in range(2):
        for field in ("grid", "coef", "scale_base", "scale_sp", "mask"):
            out[f"layers.{b}.act_fun.0.{field}"] = sd[field][b]
    return out


def stats(v):
    v = np.asarray(v, np.float64).ravel()
    return v.mean(), v.std()


def stat_close(got, want, tol=0.06):
    """The reference's assert_stat_close (tests/kan_head_init_parity.rs:56-82)."""
    if abs(want) < 1e-6:
        return abs(got) < 1e-6
    if abs(want) < 0.1:
        return abs(got - want) < 0.05
    return abs((got - want) / want) < tol


def test_reference_check_at_seed_42():
    ddr, ours = ddr_stats(), head_tensors(42)
    compared = 0
    for name, v in ours.items():
        m, s = stats(v)
        assert stat_close(m, ddr[name][0]), (name, "mean", m, ddr[name][0])
        assert stat_close(s, ddr[name][1]), (name, "std", s, ddr[name][1])
        compared += 1
    assert compared == 14
    # shapes as DDR's
    for name, v in ours.items():
        assert "x".join(map(str, v.shape)) == next(
            r["shape"] for r in csv.DictReader(open(GOLD, newline="")) if r["name"] == name)


RANDOM = {  # name -> closed-form std of the distribution
    "input.weight": math.sqrt(2.0 / F),
    "output.weight": 0.1 * math.sqrt(2.0 / (H + P)),
    "layers.0.act_fun.0.scale_base": 1.0 / math.sqrt(3.0 * H),
    "layers.1.act_fun.0.scale_base": 1.0 / math.sqrt(3.0 * H),
    "layers.0.act_fun.0.coef": None,
    "layers.1.act_fun.0.coef": None,
}


def test_ddr_sample_is_a_typical_draw_of_ours():
    ddr = ddr_stats()
    draws = [head_tensors(s) for s in SEEDS]
    for name, closed in RANDOM.items():
        ms = np.array([stats(d[name]) for d in draws])        # [seeds, (mean, std)]
        mu, sd = ms.mean(0), ms.std(0)
        want_m, want_s = ddr[name][:2]
        assert abs(want_m - mu[0]) <= 4 * sd[0], (name, "mean", want_m, mu[0], sd[0])
        assert abs(want_s - mu[1]) <= 4 * sd[1], (name, "std", want_s, mu[1], sd[1])
        # the spread is tight enough to mean something: about 1 / sqrt(2 n) for n
        # independent values (12 % for output.weight's 63), 2 % for the coefficients
        bound = {"input": 0.08, "output": 0.12, "layers": 0.05}[name.split(".")[0]]
        assert sd[1] / mu[1] < bound, (name, sd[1] / mu[1])
        if closed is not None:
            pooled = np.sqrt(np.mean([np.mean(np.square(d[name])) for d in draws]))
            n = draws[0][name].size * len(draws)
            assert abs(pooled / closed - 1) < 4 / math.sqrt(2 * n), (name, pooled, closed)
    # the coefficients' std is about 6 times the noise's (0.3 / grid / sqrt(12)):
    # a fit that only interpolated (std about the noise's) or a ridge with a
    # large penalty would fail the band above
    noise_std = 0.3 / GRID / math.sqrt(12.0)
    assert 5.0 < ddr["layers.0.act_fun.0.coef"][1] / noise_std < 7.0


def test_deterministic_tensors_exactly():
    ddr, ours = ddr_stats(), head_tensors(7)
    for b in range(2):
        g = ours[f"layers.{b}.act_fun.0.grid"]
        m, s = stats(g)
        want = ddr[f"layers.{b}.act_fun.0.grid"]
        assert abs(m) < 1e-9 and abs(s - want[1]) < 1e-6
        assert abs(g.min() - want[2]) < 1e-6 and abs(g.max() - want[3]) < 1e-6
        # every row is the same uniform extended knot vector, spacing 2 / grid
        # (held in float32: differences within an ulp of the knots)
        assert np.all(g == g[0])
        np.testing.assert_allclose(np.diff(g, axis=1), 2.0 / GRID, rtol=0, atol=3e-7)
        sp = ours[f"layers.{b}.act_fun.0.scale_sp"]
        assert np.all(np.abs(sp - ddr[f"layers.{b}.act_fun.0.scale_sp"][0]) < 1e-7)
        assert np.all(ours[f"layers.{b}.act_fun.0.mask"] == 1.0)
    for name in ("input.bias", "output.bias"):
        assert np.all(ours[name] == 0.0) and ddr[name][:2] == (0.0, 0.0)

"""The host side of the training step on the CPU: AdaDelta against the
reference's own known answers (src/training/adadelta.rs:135-230) and torch's
Adadelta (which the reference states it matches), global-norm clipping
(src/training/optimizer.rs:163-193), and the nested-gauge workload builders
against brute force on small trees."""
import math

import numpy as np
import pytest
import torch

from ddrs_amd import network as NW
from ddrs_amd import train as TR


def param(vals):
    return torch.nn.Parameter(torch.tensor(vals, dtype=torch.float32))


def test_adadelta_first_step_hand_value():
    # from zero state, rho 0.9, eps 1e-6, lr 1, g 1: delta = sqrt(1e-6)/sqrt(0.1 + 1e-6)
    p = param([1.0])
    opt = TR.AdaDelta([p])
    p.grad = torch.tensor([1.0])
    opt.step(1.0)
    want = math.sqrt(1e-6) / math.sqrt(0.1 + 1e-6)
    assert abs(p.item() - (1.0 - want)) < 1e-7
    sq, acc = opt.state[id(p)]
    assert abs(sq.item() - 0.1) < 1e-7
    assert abs(acc.item() - 0.1 * want ** 2) < 1e-12


def test_adadelta_scale_free_and_growing():
    small, huge = param([0.0]), param([0.0])
    o1, o2 = TR.AdaDelta([small]), TR.AdaDelta([huge])
    small.grad, huge.grad = torch.tensor([1.0]), torch.tensor([1000.0])
    o1.step()
    o2.step()
    assert (-huge.item()) / (-small.item()) < 1.01
    assert abs(-huge.item() - math.sqrt(10 * 1e-6)) < 1e-6
    # constant gradient: later steps exceed the first
    p = param([0.0])
    opt = TR.AdaDelta([p])
    steps = []
    for _ in range(10):
        before = p.item()
        p.grad = torch.tensor([1.0])
        opt.step()
        steps.append(before - p.item())
    assert steps[9] > steps[0]
    # a zero gradient leaves the parameter unchanged
    z = param([0.5])
    oz = TR.AdaDelta([z])
    z.grad = torch.tensor([0.0])
    oz.step()
    assert z.item() == 0.5


def test_adadelta_matches_torch_adadelta():
    rng = np.random.default_rng(0)
    init = [rng.normal(size=s).astype(np.float32) for s in ((21, 10), (21,), (2, 21))]
    ours = [param(x) for x in init]
    theirs = [param(x) for x in init]
    o1 = TR.AdaDelta(ours)
    o2 = torch.optim.Adadelta(theirs)
    for k in range(12):
        grads = [torch.tensor(rng.normal(scale=10.0 ** (k % 3 - 1), size=x.shape)
                              .astype(np.float32)) for x in init]
        for a, b, g in zip(ours, theirs, grads):
            a.grad, b.grad = g.clone(), g.clone()
        o1.step(1.0)
        o2.step()
        for a, b in zip(ours, theirs):
            np.testing.assert_allclose(a.detach().numpy(), b.detach().numpy(), rtol=2e-6,
                                       atol=1e-8)


def test_clip_grad_norm():
    a, b = param([0.0, 0.0]), param([0.0])
    a.grad, b.grad = torch.tensor([3.0, 0.0]), torch.tensor([4.0])
    norm = TR.clip_grad_norm([a, b], 1.0)
    assert norm == pytest.approx(5.0)
    np.testing.assert_allclose(a.grad.numpy(), [0.6, 0.0], rtol=1e-6)
    np.testing.assert_allclose(b.grad.numpy(), [0.8], rtol=1e-6)
    # at or under the limit: untouched
    a.grad, b.grad = torch.tensor([0.3, 0.0]), torch.tensor([0.4])
    assert TR.clip_grad_norm([a, b], 1.0) == pytest.approx(0.5)
    np.testing.assert_array_equal(a.grad.numpy(), np.float32([0.3, 0.0]))
    a.grad, b.grad = torch.zeros(2), torch.zeros(1)
    assert TR.clip_grad_norm([a, b], 1.0) == 0.0


def _brute_upstream(n, rows, cols):
    up = {i: [] for i in range(n)}
    for r, c in zip(rows, cols):
        up[int(r)].append(int(c))
    sizes = np.zeros(n, np.int64)
    members = []
    for i in range(n):
        seen, stack = {i}, [i]
        while stack:
            for c in up[stack.pop()]:
                if c not in seen:
                    seen.add(c)
                    stack.append(c)
        sizes[i] = len(seen)
        members.append(seen)
    return sizes, members


def test_nested_gauges_and_union_brute_force():
    n = 600
    rows, cols = NW.gen_tree(n, 16, 77)
    sizes, members = _brute_upstream(n, rows, cols)
    np.testing.assert_array_equal(TR.upstream_sizes(n, rows, cols), sizes)
    g = TR.nested_gauges(n, rows, cols, 25, 3, 300, seed=5)
    assert g.size == 25 and np.unique(g).size == 25 and (np.diff(g) > 0).all()
    # each gauge's basin is the unused reach nearest to some log-uniform draw
    assert sizes[g].min() >= 1 and sizes[g].max() <= n
    ur, uc = TR.upstream_union(n, rows, cols, g)
    covered = set().union(*[members[int(x)] for x in g])
    want = [(int(r), int(c)) for r, c in zip(rows, cols) if int(r) in covered]
    assert list(zip(ur.tolist(), uc.tolist())) == want
    # compression keeps a lower-triangular tree on the covered reaches
    active, cr, cc, cg = NW.compress(ur, uc, g)
    assert set(active.tolist()) == covered
    np.testing.assert_array_equal(active[cg], g)
    assert (cr > cc).all()


def test_nested_gauges_draws_follow_targets():
    n = 20000
    rows, cols = NW.gen_tree(n, 256, 3)
    g = TR.nested_gauges(n, rows, cols, 200, 10, 5000, seed=9)
    s = TR.upstream_sizes(n, rows, cols)[g]
    rs = np.random.default_rng(9)
    targets = np.sort(np.exp(rs.uniform(math.log(10), math.log(5000), 200)))
    # the sorted basin sizes track the sorted log-uniform draws
    ratio = np.sort(s) / targets
    assert np.median(np.abs(np.log(ratio))) < 0.1


@pytest.mark.parametrize("ddr_match", [False, True])
def test_oracle_training_chain_matches_fp64_finite_differences(ddr_match):
    """Pins the chain the GPU training test compares against: its analytic
    weight gradients (KAN backward of the routing adjoint of the loss
    gradient, through the denormalize slopes) agree with float64 central
    differences of its own loss, for each gauge's own reach and for the
    ddr_match convention (physics, and a prediction summed over the gauge's
    upstream reaches)."""
    from ddrs_amd import kan as K
    from tests import train_case as TC
    b = TC.Batch(n_tree=1500, gauges=8, T=120, seed=3, ddr_match=ddr_match)
    if ddr_match:   # some gauge sums several upstream reaches
        assert np.bincount(b.outflow[2]).max() > 1
    mk = lambda s: K.KanHeadConfig([f"a{i}" for i in range(TC.F)], TC.NAMES, s).init("cpu")
    b.observe(TC.oracle_params(mk(99)), 3)
    head = mk(42)
    p = TC.oracle_params(head)
    loss, nv, g = b.oracle_loss_and_grads(p, 3, dtype=np.float64)
    assert nv == (b.gauges.size - 1) * ((b.T - 24) // 24 - b.warmup)
    picks = [("output_bias", (0,)), ("output_bias", (1,)), ("input_weight", (4, 2)),
             ("input_bias", (7,)), ("coef", (1, 3, 5, 2)), ("scale_base", (0, 2, 7)),
             ("scale_sp", (1, 11, 4))]
    for name, idx in picks:
        def at(delta):
            q = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in p.items()}
            q["blocks"] = [{k: v.copy() for k, v in blk.items()} for blk in p["blocks"]]
            if name in TC.BLOCK:
                q["blocks"][idx[0]][name][idx[1:]] += delta
            else:
                q[name][idx] += delta
            return b.oracle_loss(q, 3, dtype=np.float64)[0]
        h = 1e-5
        fd = (at(h) - at(-h)) / (2 * h)
        scale = float(np.abs(g[name]).max())
        assert abs(g[name][idx] - fd) <= 1e-5 * scale + 1e-6 * abs(fd), (name, g[name][idx], fd)

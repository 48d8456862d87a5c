"""Parity at the measured workloads' full sizes: the HIP path through the C ABI
against the oracle on the benchmark's own networks, forcing and gauges
(ddrs_amd.workloads). Needs an MI355X.

c2  the whole 350,000-reach x 720-step window: discharge at every reach-step
    (a dense forward) and at the 64 gauges, the gradients w.r.t. n and
    q_spatial of the gauge MSE for every reach, and FP64 central differences
    on 64 seeded reaches.
c3 / c4
    forests of independent basins. No edge joins two basins, so a basin
    routed alone is routed exactly as inside the forest (the oracle's work per
    reach reads only the reach's upstream reaches, in CSR order, which the
    relabelling keeps). From one GPU run of the whole forest, with the
    workload's checkpoint segments, whole basins are compared at full length:
    the largest (c4: its first 720 hours, with the loss's weight on that basin
    zero after them, which makes its gradients those of a 720-hour window)
    and a seeded sample of the rest.
c5  the training micro-batch: the routing of a sample of its trees (discharge
    at every reach, gradients of the batch NSE through the device's own
    dL/dQ at the gauges), the loss stage on the device's gauge discharge, and
    the KAN backward of the device's per-reach gradients over all reaches.

Three oracles take part (oracle/scale_check.py): the f32 restatement of the
reference (glibc's powf and logf, as the reference's libm calls), the same
float arithmetic with the device's correctly rounded powers and logarithms
("device-math"), and FP64. glibc's powf rounds about 0.06% of the path's powers
away from the correctly rounded result, and at full size the recurrence
amplifies single roundings past 1e-4 at a few reach-steps (with Cunge X most
reaches have c1 < 0; on c2 the f32 oracle is up to 12% from FP64 at some).
There two f32 evaluations that round one power differently part by more than
1e-4 while both stay as far from FP64. The device-math oracle separates that
from the kernels' own arithmetic.

Bars. Discharge: within 1e-5 relative of the device-math oracle at every
compared reach-step (bit-identical is expected: the fraction is reported);
within 1e-4 of the f32 oracle at 99.999% of them, and nowhere farther from it
than twice the spread between the f32 and device-math oracles; an FP64 error
at most 3x the f32 oracle's own. Gradients: vectors within 1e-4 max|g| of the
device-math oracle (the adjoint's gradient-only divisions use the hardware
reciprocal) and within 1e-3 max|g| of the f32 oracle, or twice the spread
between the two oracles where libm's rounding moves that one further, with an
FP64 error at most 3x the f32 oracle's own. Beside the norm bar, the element-wise
relative error on the well-conditioned entries (|g| >= 1e-2 max|g|) is
reported, and against the device-math oracle it is held within 5e-3 (a
systematic relative error on small but clean gradients fails it). The finite differences (oracle/scale_check.py: the step chosen per
reach from five, by the agreement of adjacent steps): on every reach they
resolve (two steps agreeing within 1e-3; at least 40 of 64) the f64 adjoint
within 2e-3 of them, which pins the differences, and each GPU gradient within
max(1e-3, 3x the f32 oracle's error) of them. A reach whose gradient vanishes
below 1e-6 of the largest (thousands of damping links above the nearest gauge)
is not resolved and is reported relative to the largest.
"""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor, as_completed

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import scale_check as SC

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on the CPU too, skipped there
    pytest.skip("needs a GPU", allow_module_level=True)

from ddrs_amd import _native as NAT  # noqa: E402
from ddrs_amd import workloads as W  # noqa: E402
from ddrs_amd.sparse import Network  # noqa: E402

DEV = "cuda:0"
OUT = os.environ.get("DDRS_SCALE_REPORT")   # optional JSON report of every figure


@pytest.fixture(autouse=True)
def _release_device_memory():
    """The workloads fill most of the device: free each test's buffers and
    workspaces before the next one starts."""
    yield
    import gc
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def report(name, figures):
    print(f"{name}: {json.dumps(figures, default=float)}", flush=True)
    if OUT:
        rep = {}
        if os.path.exists(OUT):
            with open(OUT) as fh:
                rep = json.load(fh)
        rep[name] = figures
        with open(OUT, "w") as fh:
            json.dump(rep, fh, indent=1, default=float)


def assert_discharge(label, e):
    """The bars of the module docstring on discharge_errors' figures."""
    assert e["max_rel_err_vs_device_math_oracle"] <= 1e-5, (label, e)
    assert e["frac_within_1e-4"] >= 0.99999, (label, e)
    assert e["max_rel_err_vs_f32_oracle"] <= (
        2.0 * e["f32_oracle_vs_device_math_oracle_max_rel"] + 1e-5), (label, e)
    assert e["max_rel_err_vs_fp64"] <= 3.0 * e["f32_oracle_max_rel_err_vs_fp64"] + 1e-6, (label, e)


def assert_grad(label, e):
    assert e["norm_err_vs_f32_oracle"] <= max(
        1e-3, 2.0 * e["f32_oracle_vs_device_math_oracle_norm"] + 1e-4), (label, e)
    assert e["norm_err_vs_device_math_oracle"] <= 1e-4, (label, e)
    # element-wise on the well-conditioned entries (|g| >= 1e-2 max|g|): the
    # adjoint's gradient-only divisions use the hardware reciprocal, and the
    # per-step terms of an entry sum over thousands of steps (measured at most
    # 1.1e-3 over every compared reach of c2-c5)
    assert e["elem_max_rel_err_vs_device_math_oracle"] <= 5e-3, (label, e)
    assert e["norm_err_vs_fp64"] <= 3.0 * e["f32_oracle_norm_err_vs_fp64"] + 1e-6, (label, e)


@pytest.mark.timeout(1200)
def test_c2_full_size_parity():
    """c2 (350,000 reaches x 720 steps): the benchmark step's forward and
    adjoint against the oracle at every reach and step, and FP64 finite
    differences on 64 seeded reaches."""
    t0 = time.time()
    wl = W.build("c2")
    R, T = wl.R, wl.T
    say(f"c2 {R} x {T}: building")
    qd = wl.forcing_device()
    qp = qd.cpu().numpy()
    # the host generator (the CPU baseline's forcing) against the device's
    qh = wl.forcing_host()
    ulp = np.abs(qh.view(np.int32).astype(np.int64) - qp.view(np.int32).astype(np.int64))
    gen = {"identical_frac": float((ulp == 0).mean()), "max_ulp": int(ulp.max())}
    del qh, ulp
    say(f"forcing host vs device: {gen}")
    assert gen["max_ulp"] <= 1
    net = Network(wl.adjacency())
    gw = W.GaugedWindow(net, wl, qd)
    Q = gw.forward().clone()
    obs = 0.9 * Q
    gw.step(obs)
    Qd = gw.dense_forward()
    torch.cuda.synchronize()
    gw.check()
    Qg, Qdense = gw.Q.cpu().numpy(), Qd.cpu().numpy()
    gQ = gw.gQ.cpu().numpy()
    got = {"n": gw.gn.cpu().numpy(), "q_spatial": gw.gqsp.cpu().numpy()}
    q0 = gw.q0.cpu().numpy()
    Q1 = Q.cpu().numpy()
    gw.close()
    del Qd, gw, qd, Q, obs
    say(f"gpu done ({time.time() - t0:.0f}s)")
    assert np.array_equal(Qg, Qdense[:, wl.gauges]), "gauged readout != dense columns"
    np.testing.assert_array_equal(Qg, Q1)   # a second forward gives the same bits
    P = O.make_params()
    pat = O.pattern_from_coo(R, wl.rows, wl.cols)
    np.testing.assert_array_equal(q0, O.hotstart(P, pat, qp[0]))
    at = wl.attrs
    p = np.full(R, 21.0, np.float32)
    gQd = np.zeros((T, R), np.float32)
    gQd[:, wl.gauges] = gQ
    with ThreadPoolExecutor(SC.threads()) as pool:
        r = SC.route_pair(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp,
                          gQd, q0, want_x64=True, pool=pool)
        say(f"oracle f32 + f64 done ({time.time() - t0:.0f}s)")
        figs = {"forcing_host_vs_device": gen,
                "discharge_dense": SC.discharge_errors(Qdense, r),
                "discharge_gauges": SC.discharge_errors(Qg, r, wl.gauges)}
        del Qdense
        for k in SC.GRAD_KEYS:
            figs[f"grad_{k}"] = SC.grad_errors(got[k], r, k)
        rng = np.random.default_rng(4102)
        reaches = np.sort(rng.choice(R, 64, replace=False))
        params = {"n": at["n"], "q_spatial": at["q_spatial"], "p_spatial": p}
        fd = SC.fd_check(P, pat, params, at["length"], at["slope"], qp, r["f64"]["Q"], r["f64"]["X"],
                         gQ, wl.gauges, reaches, pool=pool)
    figs["fd_64_reaches"] = SC.fd_errors(fd, reaches, got, r)
    figs["seconds"] = time.time() - t0
    report("c2", figs)
    assert_discharge("c2 dense", figs["discharge_dense"])
    assert_discharge("c2 gauges", figs["discharge_gauges"])
    for k in SC.GRAD_KEYS:
        assert_grad(f"c2 {k}", figs[f"grad_{k}"])
        e = figs["fd_64_reaches"][k]
        assert e["resolved"] >= 40, (k, e["resolved"])
        assert e["max_rel_err_fp64_adjoint"] <= 2e-3, (k, e["max_rel_err_fp64_adjoint"])
        for g, o, ok in zip(e["gpu_rel_err_per_reach"], e["f32_oracle_rel_err_per_reach"],
                            e["resolved_mask"]):
            if ok:
                assert g <= max(1e-3, 3.0 * o), (k, g, o)


def _basin_parity(wl, b, lo, s, T_b, Qg, gQ, gn, gqsp, q0, qp):
    """Basin b (reaches [lo, lo + size)) routed alone by the oracle for T_b
    steps against the forest run's values (readout slots [s, s + size))."""
    size = int(wl.sizes[b])
    rows, cols = wl.sub_basin(b)
    P = O.make_params()
    pat = O.pattern_from_coo(size, rows, cols)
    at = {k: v[lo:lo + size] for k, v in wl.attrs.items()}
    p = np.full(size, 21.0, np.float32)
    q0b = q0[lo:lo + size]
    np.testing.assert_array_equal(q0b, O.hotstart(P, pat, qp[0]), err_msg=f"basin {b} q0")
    r = SC.route_pair(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp, gQ, q0b)
    figs = {"reaches": size, "steps": T_b, "discharge": SC.discharge_errors(Qg, r)}
    got = {"n": gn[lo:lo + size], "q_spatial": gqsp[lo:lo + size]}
    for k in SC.GRAD_KEYS:
        figs[f"grad_{k}"] = SC.grad_errors(got[k], r, k)
    return figs


def _forest_parity(config, n_random, mid_sizes, cut=None):
    """One GPU run of the whole forest (the benchmark's MSE at the loss
    gauges, the workload's checkpoint segments), whole basins read out at
    every reach and compared with the oracle routing each basin alone. cut:
    the largest basin is compared over its first `cut` steps, its loss weight
    zero after them."""
    t0 = time.time()
    wl = W.build(config)
    R, T = wl.R, wl.T
    nb = wl.sizes.size
    big = int(np.argmax(wl.sizes))
    rng = np.random.default_rng(4100 + int(config[1:]))
    rest = np.setdiff1d(np.arange(nb), [big])
    sample = {big} | set(int(b) for b in rng.choice(rest, n_random, replace=False))
    for m in mid_sizes:   # the basins nearest in size to each of mid_sizes
        sample.add(int(rest[np.argmin(np.abs(wl.sizes[rest] - m))]))
    sample = sorted(sample)
    readout = np.unique(np.concatenate([wl.gauges] + [wl.basin_reaches(b) for b in sample]))
    readout = readout.astype(np.int32)
    loss_slots = torch.as_tensor(np.searchsorted(readout, wl.gauges), device=DEV)
    say(f"{config}: {R} reaches in {nb} basins x {T} steps, checkpoint {wl.ckpt}; "
        f"{len(sample)} basins compared ({wl.sizes[sample].sum()} reaches, largest "
        f"{wl.sizes[big]}), {readout.size} reaches read out")
    qd = wl.forcing_device()
    net = Network(wl.adjacency())
    gw = W.GaugedWindow(net, wl, qd, gauges=readout, window=wl.ckpt)
    Q = gw.forward()
    with torch.no_grad():
        Ql = Q.index_select(1, loss_slots)
        gw.gQ.zero_()
        gw.gQ.index_copy_(1, loss_slots, (Ql - 0.9 * Ql) * (2.0 / (T * wl.gauges.size)))
        del Ql
        big_out = int(wl.offsets[big] + wl.sizes[big] - 1)
        if cut is not None:
            gw.gQ[cut:, int(np.searchsorted(readout, big_out))] = 0.0
    gw.adjoint()
    torch.cuda.synchronize()
    gw.check()
    say(f"{config}: gpu done ({time.time() - t0:.0f}s)")
    gn, gqsp, q0 = (v.cpu().numpy() for v in (gw.gn, gw.gqsp, gw.q0))
    jobs = []
    for b in sample:
        lo, size = int(wl.offsets[b]), int(wl.sizes[b])
        s = int(np.searchsorted(readout, lo))
        T_b = cut if (b == big and cut is not None) else T
        take = lambda a: a[:T_b, s:s + size].contiguous().cpu().numpy()
        jobs.append((b, lo, s, T_b, take(gw.Q), take(gw.gQ),
                     qd[:T_b, lo:lo + size].contiguous().cpu().numpy()))
    gw.close()
    del gw, qd, Q
    torch.cuda.empty_cache()
    with ThreadPoolExecutor(max(1, SC.threads() // 2)) as pool:
        futs = {pool.submit(_basin_parity, wl, b, lo, s, T_b, Qb, gQb, gn, gqsp, q0, qpb): b
                for b, lo, s, T_b, Qb, gQb, qpb in jobs}
        per = {}
        for f in as_completed(futs):
            b = futs[f]
            per[b] = f.result()
            say(f"{config}: basin {b} ({wl.sizes[b]} reaches) compared, discharge max rel err "
                f"{per[b]['discharge']['max_rel_err_vs_f32_oracle']:.2e} ({time.time() - t0:.0f}s)")
    worst = {"discharge_max_rel_err_vs_f32_oracle":
             max(f["discharge"]["max_rel_err_vs_f32_oracle"] for f in per.values()),
             "discharge_max_rel_err_vs_fp64":
             max(f["discharge"]["max_rel_err_vs_fp64"] for f in per.values())}
    for k in SC.GRAD_KEYS:
        worst[f"grad_{k}_norm_err_vs_f32_oracle"] = max(
            f[f"grad_{k}"]["norm_err_vs_f32_oracle"] for f in per.values())
        worst[f"grad_{k}_elem_max_rel_err_vs_f32_oracle"] = max(
            f[f"grad_{k}"]["elem_max_rel_err_vs_f32_oracle"] for f in per.values())
    figs = {"reaches": R, "basins": int(nb), "steps": T, "checkpoint": wl.ckpt,
            "largest_basin": {"basin": big, **per[big]}, "worst_over_basins": worst,
            "basins_compared": len(per), "reaches_compared": int(wl.sizes[sample].sum()),
            "per_basin": {str(b): f for b, f in per.items()}, "seconds": time.time() - t0}
    report(config, figs)
    for b, f in per.items():
        assert_discharge(f"{config} basin {b}", f["discharge"])
        for k in SC.GRAD_KEYS:
            assert_grad(f"{config} basin {b} {k}", f[f"grad_{k}"])


@pytest.mark.timeout(1200)
def test_c3_forest_sampled_basins():
    """c3: 512 basins (about 3 M reaches) x 2,160 steps in checkpoint segments
    of 256; the largest basin and 20 seeded others, plus those nearest 1,000
    and 10,000 reaches, compared whole at full length."""
    _forest_parity("c3", 20, (1000, 10000))


@pytest.mark.timeout(1200)
def test_c4_forest_sampled_basins():
    """c4: 200,000 basins (2.94 M reaches; Zipf(1.8) sizes capped at 300,000,
    then scaled to the total, so the largest holds 111,430) x 8,760 steps in
    checkpoint segments of 168; the largest basin over its first 720 steps, 20
    seeded others and those nearest 1,000 and 10,000 reaches at full length."""
    _forest_parity("c4", 20, (1000, 10000), cut=720)


def _tree_parity(wl, reach_ids, T, Qg, gQ, gn, gqsp, nq, q0, qp, grads=True):
    """One tree of the c5 batch (positions reach_ids, ascending) routed alone by
    the oracle with the device's n and q_spatial, against the batch run (the
    discharge only, without grads)."""
    R = reach_ids.size
    loc = np.full(wl.R, -1, np.int64)
    loc[reach_ids] = np.arange(R)
    keep = loc[wl.cols] >= 0
    rows, cols = loc[wl.rows[keep]].astype(np.int32), loc[wl.cols[keep]].astype(np.int32)
    P = O.make_params()
    pat = O.pattern_from_coo(R, rows, cols)
    at = {k: v[reach_ids] for k, v in wl.attrs.items()}
    n, qsp = nq[0][reach_ids], nq[1][reach_ids]
    p = np.full(R, 21.0, np.float32)
    q0t = q0[reach_ids]
    np.testing.assert_array_equal(q0t, O.hotstart(P, pat, qp[0]))
    if not grads:
        r = SC.route_fwd_set(P, pat, n, qsp, p, at["length"], at["slope"], qp, q0t)
        return {"reaches": int(R), "steps": T, "discharge": SC.discharge_errors(Qg, r)}
    r = SC.route_pair(P, pat, n, qsp, p, at["length"], at["slope"], qp, gQ, q0t)
    figs = {"reaches": int(R), "steps": T, "discharge": SC.discharge_errors(Qg, r)}
    got = {"n": gn[reach_ids], "q_spatial": gqsp[reach_ids]}
    for k in SC.GRAD_KEYS:
        figs[f"grad_{k}"] = SC.grad_errors(got[k], r, k)
    return figs


@pytest.mark.timeout(1200)
def test_c5_training_batch_sampled_trees():
    """c5: the training micro-batch (285,985 reaches in 62 trees, 2,000 gauges,
    8,760 steps): the device's loss stage; at full length the discharge of the
    largest tree (60,997 reaches) and the discharge and gradients of the trees
    nearest 5,000 and 12,000 reaches and 8 seeded others; and the KAN backward
    of the device's per-reach routing gradients over every reach."""
    from oracle import kan_oracle as ko
    from oracle import loss_oracle as lo
    from tests import train_case as TC
    from ddrs_amd import train as TR
    t0 = time.time()
    cfg = W.CONFIGS["c5"]
    wl = W.build_c5()
    R, T = wl.R, wl.T
    G = wl.gauges.size
    qd = wl.forcing_device()
    net = Network(wl.adjacency())
    batch = W.c5_batch(wl, qd, net)
    head = W.kan_head(42)
    tr = TR.Trainer(head, tau=cfg["tau"], warmup=cfg["warmup"], eps=cfg["eps"], network=net)
    trace = {}
    loss, n_valid = tr.loss(batch, trace)
    loss.backward()
    torch.cuda.synchronize()
    say(f"c5: {R} reaches, {G} gauges, {T} steps; gpu step done ({time.time() - t0:.0f}s)")
    Qg = trace["Qg"].detach().cpu().numpy()          # [T, G]
    gQg = trace["Qg"].grad.cpu().numpy()
    nq = (trace["n"].detach().cpu().numpy(), trace["q_spatial"].detach().cpu().numpy())
    gn, gqsp = trace["n"].grad.cpu().numpy(), trace["q_spatial"].grad.cpu().numpy()
    q0 = trace["q0"].detach().cpu().numpy()
    grads = {name: getattr(head, name).grad.double().cpu().numpy() for name in TC.DENSE + TC.BLOCK}
    obs, sigma = batch.observations.cpu().numpy(), batch.gauge_obs_std.cpu().numpy()
    figs = {"reaches": R, "gauges": G, "steps": T}
    # the loss stage on the device's gauge discharge
    ol, onv, og = lo.gauge_nse_loss(Qg.T.astype(np.float64), np.arange(G), np.arange(G), G,
                                    cfg["tau"], cfg["warmup"], obs, sigma, cfg["eps"])
    figs["loss"] = {"device": float(loss), "oracle": ol, "rel_err": abs(float(loss) - ol) / abs(ol),
                    "n_valid": int(n_valid), "oracle_n_valid": onv,
                    "dLdQ_norm_err": float(np.abs(gQg - og.T).max() / np.abs(og).max())}
    # the trees: the largest, those nearest 5,000 and 12,000 reaches, 8 seeded others
    tree = W.trees(R, wl.rows, wl.cols)
    outs, size = np.unique(tree, return_counts=True)
    order = np.argsort(size)[::-1]
    pick = {int(order[0])}
    for m in (5000, 12000):
        pick.add(int(np.argmin(np.abs(size - m))))
    rng = np.random.default_rng(4105)
    pick |= set(int(i) for i in rng.choice(np.setdiff1d(np.arange(outs.size), list(pick)), 8,
                                           replace=False))
    members = {i: np.flatnonzero(tree == outs[i]) for i in sorted(pick)}
    readout = np.concatenate(list(members.values())).astype(np.int32)
    readout.sort()
    # a second workspace beside the trainer's one-segment one: checkpoint
    # segments of 1,024 steps keep it small
    gw = W.GaugedWindow(net, wl, qd, gauges=readout, window=1024, n=trace["n"],
                        q_spatial=trace["q_spatial"])
    gw.forward()
    torch.cuda.synchronize()
    gw.check()
    Qr = gw.Q
    slot = {int(g): k for k, g in enumerate(wl.gauges)}
    jobs = []
    for i, ids in members.items():
        s = torch.as_tensor(np.searchsorted(readout, ids), device=DEV)
        gQt = np.zeros((T, ids.size), np.float32)
        for k, r in enumerate(ids):
            if int(r) in slot:
                gQt[:, k] = gQg[:, slot[int(r)]]
        jobs.append((i, ids, Qr.index_select(1, s).cpu().numpy(), gQt,
                     qd.index_select(1, torch.as_tensor(ids, device=DEV)).cpu().numpy()))
    gw.close()
    del gw, Qr
    say(f"c5: {len(jobs)} trees ({sum(j[1].size for j in jobs)} reaches) to the oracle "
        f"({time.time() - t0:.0f}s)")
    with ThreadPoolExecutor(max(1, SC.threads() // 2)) as pool:
        # the largest tree (61k reaches x 8,760 steps) is compared on its
        # discharge: its adjoint alone would take the oracle several minutes
        futs = {pool.submit(_tree_parity, wl, ids, T, Qt, gQt, gn, gqsp, nq, q0, qpt,
                            i != int(order[0])): i
                for i, ids, Qt, gQt, qpt in jobs}
        # the KAN backward over every reach in chunks (its gradients are sums
        # over reaches), from the device's dL/d(n, q_spatial)
        p = TC.oracle_params(head)
        x = batch.attributes.cpu().numpy()
        gout = np.stack([gn.astype(np.float64) * (TC.N_RANGE[1] - TC.N_RANGE[0]),
                         gqsp.astype(np.float64) * (TC.QSP_RANGE[1] - TC.QSP_RANGE[0])], axis=1)

        def kan_chunk(lo_, hi_):
            _, cache = ko.kan_head_fwd(x[lo_:hi_], p, head.k)
            g = ko.kan_head_bwd(gout[lo_:hi_], p, cache)
            out = {k: g[k] for k in TC.DENSE}
            for k in TC.BLOCK:
                out[k] = np.stack([b[k] for b in g["blocks"]])
            return out

        step = 32768
        kf = [pool.submit(kan_chunk, a, min(a + step, R)) for a in range(0, R, step)]
        kan = {}
        for f in kf:
            for k, v in f.result().items():
                kan[k] = kan.get(k, 0.0) + v
        say(f"c5: KAN backward oracle done ({time.time() - t0:.0f}s)")
        per = {}
        for f in as_completed(futs):
            i = futs[f]
            per[i] = f.result()
            say(f"c5: tree of {size[i]} reaches compared, discharge max rel err "
                f"{per[i]['discharge']['max_rel_err_vs_f32_oracle']:.2e} ({time.time() - t0:.0f}s)")
    figs["kan_backward_norm_err"] = {k: float(np.abs(grads[k] - kan[k]).max() /
                                              np.abs(kan[k]).max()) for k in kan}
    figs["largest_tree"] = per[int(order[0])]
    figs["trees_compared"] = len(per)
    figs["reaches_compared"] = int(sum(f["reaches"] for f in per.values()))
    figs["per_tree"] = {str(int(size[i])): f for i, f in per.items()}
    figs["seconds"] = time.time() - t0
    report("c5", figs)
    assert figs["loss"]["n_valid"] == onv
    assert figs["loss"]["rel_err"] <= 1e-4 and figs["loss"]["dLdQ_norm_err"] <= 1e-4
    for i, f in per.items():
        assert_discharge(f"c5 tree {i}", f["discharge"])
        for k in SC.GRAD_KEYS:
            if f"grad_{k}" in f:
                assert_grad(f"c5 tree {i} {k}", f[f"grad_{k}"])
    for k, e in figs["kan_backward_norm_err"].items():
        assert e <= 1e-4, (k, e)

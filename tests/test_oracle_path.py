"""The oracle's downstream-path re-route (orc_route_path), used for FP64 central
finite differences on workload-sized networks, pinned against full re-routes.

A change of reach r's parameters moves the discharge of r and of the reaches
downstream of it and of no other reach. The path re-route recomputes only those
reaches, reading every other reach from the unchanged run, with the arithmetic
of the full step. So it must give the full re-route's values bit for bit on the
path, in float and in double, and the full re-route must leave every reach off
the path bit-identical. Each of three planted mistakes in the re-route fails the
first test: the solve reading an upstream reach's post-clamp discharge instead
of its pre-clamp solve; the solve reading the unchanged run's value for the
path's own upstream reach; the SpMV reading the unchanged run's discharge for
it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ddrs_amd import network as NW


def _case(R, T, seed, P):
    rows, cols = NW.gen_tree(R, 32, 100 + seed)
    at = NW.gen_attributes(R, 200 + seed)
    qp = NW.gen_forcing(R, T, 300 + seed, storms=True)
    pat = O.pattern_from_coo(R, rows, cols)
    q0 = O.hotstart(P, pat, qp[0])
    return pat, at, qp, q0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("variant", ["default", "positivity", "ddr_match"])
def test_path_reroute_equals_full_reroute(dtype, variant):
    P = {"default": O.make_params(), "positivity": O.make_params(enforce_positivity=True),
         "ddr_match": O.make_params(ddr_match=True)}[variant]
    R, T = 400, 40
    pat, at, qp, q0 = _case(R, T, 1, P)
    p = np.full(R, 21.0)
    base = dict(n=at["n"].astype(dtype), qsp=at["q_spatial"].astype(dtype), p=p.astype(dtype))
    Qb, Xb = O.route_fwd(P, pat, base["n"], base["qsp"], base["p"], at["length"], at["slope"],
                         qp, q0=q0, dtype=dtype, want_x=True)
    rng = np.random.default_rng(5)
    for r in rng.choice(R, 6, replace=False):
        path = O.downstream_path(pat, r)
        assert path[-1] == R - 1
        # unchanged parameters: the path re-route reproduces the base run
        same = O.route_path(P, pat, base["n"], base["qsp"], base["p"], at["length"],
                            at["slope"], qp, Qb, Xb, path, dtype=dtype)
        assert np.array_equal(same, Qb[:, path])
        for key in ("n", "qsp", "p"):
            v = {k: x.copy() for k, x in base.items()}
            v[key][r] = v[key][r] * dtype(1.01)
            full = O.route_fwd(P, pat, v["n"], v["qsp"], v["p"], at["length"], at["slope"], qp,
                               q0=q0, dtype=dtype)
            part = O.route_path(P, pat, v["n"], v["qsp"], v["p"], at["length"], at["slope"],
                                qp, Qb, Xb, path, dtype=dtype)
            assert np.array_equal(full[:, path], part), (variant, r, key)
            off = np.setdiff1d(np.arange(R), path)
            assert np.array_equal(full[:, off], Qb[:, off]), (variant, r, key)
            assert not np.array_equal(part, Qb[:, path]), "the change must move the path"


def test_route_adj_returns_forward():
    P = O.make_params()
    R, T = 300, 20
    pat, at, qp, q0 = _case(R, T, 2, P)
    p = np.full(R, 21.0, np.float32)
    gQ = np.random.default_rng(1).normal(size=(T, R)).astype(np.float32)
    for dt in (np.float32, np.float64):
        Q, X = O.route_fwd(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp,
                           q0=q0, dtype=dt, want_x=True)
        g = O.route_adj(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp, gQ,
                        q0=q0, dtype=dt, want_q=True, want_x=True)
        g0 = O.route_adj(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp, gQ,
                         q0=q0, dtype=dt)
        assert np.array_equal(g["Q"], Q) and np.array_equal(g["X"], X)
        for k in ("n", "q_spatial", "p_spatial", "q_prime"):
            assert np.array_equal(g[k], g0[k])


def test_path_fd_matches_full_fd_and_adjoint():
    """Central differences through the path re-route equal those through full
    re-routes (the same numbers), and agree with the f64 adjoint."""
    P = O.make_params()
    R, T = 500, 30
    pat, at, qp, q0 = _case(R, T, 3, P)
    p = np.full(R, 21.0)
    rng = np.random.default_rng(9)
    gauges = np.sort(np.concatenate([[R - 1], rng.choice(R - 1, 7, replace=False)]))
    obs = rng.uniform(0.5, 2.0, (T, gauges.size))
    n64, q64 = at["n"].astype(np.float64), at["q_spatial"].astype(np.float64)
    Qb, Xb = O.route_fwd(P, pat, n64, q64, p, at["length"], at["slope"], qp,
                         q0=q0.astype(np.float64), dtype=np.float64, want_x=True)
    gQ = np.zeros((T, R))
    gQ[:, gauges] = 2.0 * (Qb[:, gauges] - obs)
    g = O.route_adj(P, pat, n64, q64, p, at["length"], at["slope"], qp, gQ,
                    q0=q0.astype(np.float64), dtype=np.float64)

    def loss(Q):
        return float(((Q[:, gauges] - obs) ** 2).sum())

    worst = 0.0
    for r in rng.choice(R, 5, replace=False):
        path = O.downstream_path(pat, r)
        on = np.isin(path, gauges)
        gi = np.searchsorted(gauges, path[on])
        for key, base in (("n", n64), ("q_spatial", q64)):
            # the recurrence amplifies rounding about a thousandfold even in
            # double, so a relative step of 1e-6 leaves ~1e-3 of noise on small
            # gradients; at 1e-4 truncation and noise both stay below ~1e-5
            h = 1e-4 * abs(base[r])
            vals = []
            for s in (+1, -1):
                v = base.copy()
                v[r] += s * h
                nn, qq = (v, q64) if key == "n" else (n64, v)
                full = O.route_fwd(P, pat, nn, qq, p, at["length"], at["slope"], qp,
                                   q0=q0.astype(np.float64), dtype=np.float64)
                part = O.route_path(P, pat, nn, qq, p, at["length"], at["slope"], qp, Qb, Xb,
                                    path)
                assert np.array_equal(full[:, path], part)
                vals.append((loss(full), part[:, on]))
            fd_full = (vals[0][0] - vals[1][0]) / (2 * h)
            qpl, qmi = vals[0][1], vals[1][1]
            fd_path = float(((qpl - qmi) * (qpl + qmi - 2 * obs[:, gi])).sum()) / (2 * h)
            # the difference of two whole-loss sums loses about eps |L| / h to
            # cancellation; the path form takes the difference term by term
            roundoff = 64 * np.finfo(np.float64).eps * abs(vals[0][0]) / h
            assert abs(fd_path - fd_full) <= roundoff, (fd_path, fd_full, roundoff)
            worst = max(worst, abs(g[key][r] - fd_path) / abs(fd_path))
    assert worst < 1e-4, worst


def test_pattern_build_is_thread_safe():
    """The workload-size checks build many basins' patterns in parallel
    threads: concurrent builds must equal serial ones (the build keeps no
    shared state)."""
    from concurrent.futures import ThreadPoolExecutor
    cases = []
    for seed in range(16):
        R = 40000 + 1009 * seed
        rows, cols = NW.gen_tree(R, 8 + seed, 900 + seed)
        cases.append((R, rows, cols))
    serial = [O.pattern_from_coo(R, r, c) for R, r, c in cases]
    with ThreadPoolExecutor(8) as pool:
        for _ in range(3):
            par = list(pool.map(lambda a: O.pattern_from_coo(*a), cases))
            for a, b in zip(serial, par):
                for f in ("crow", "col", "row_for_nnz", "adj_values", "trans_col", "trans_to_orig"):
                    assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_threaded_oracle_is_bit_identical(dtype):
    """The oracle's per-reach loops on several threads (the CPU baseline's
    parallel leg) give the serial oracle's bits: each reach's values are
    computed on their own, and the solves and scatters stay serial."""
    R, T = 3000, 30
    rows, cols = NW.gen_tree(R, 64, 1301)
    at = NW.gen_attributes(R, 2301)
    qp = NW.gen_forcing(R, T, 3301, storms=True)
    P = O.make_params()
    pat = O.pattern_from_coo(R, rows, cols)
    p = np.full(R, 21.0, np.float32)
    gQ = np.random.default_rng(5).normal(0, 1, (T, R)).astype(np.float32)
    runs = []
    for threads in (1, 4):
        O.set_threads(threads)
        try:
            runs.append(O.route_adj(P, pat, at["n"], at["q_spatial"], p, at["length"],
                                    at["slope"], qp, gQ, dtype=dtype))
        finally:
            O.set_threads(1)
    for k in runs[0]:
        np.testing.assert_array_equal(runs[1][k], runs[0][k], err_msg=k)

"""Pins the oracle's hand-written adjoint against FP64 central finite
differences (what the mathematics fixes) on the reference's gradcheck fixtures
and on routed windows. CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.nets import POS_LENGTH, POS_QT_GRAD, SP8_INPUTS, chain, mock_params, random_tree, stress_params

PARENTS = ["n", "qsp", "p", "q_t", "qp_t"]
GRAD_KEY = dict(n="n", qsp="q_spatial", p="p_spatial", q_t="q_t", qp_t="q_prime_t")


def _step_loss(P, pat, x, L, S, w):
    qn, _, _ = O.step_fwd(P, pat, x["n"], x["qsp"], x["p"], x["q_t"], x["qp_t"], L, S,
                          dtype=np.float64)
    return float(np.dot(w, qn))


def _check_step(P, pat, x, L, S, rel_tol):
    R = pat.n
    w = 1.0 / (1.0 + np.arange(R))
    qn, saved, a = O.step_fwd(P, pat, x["n"], x["qsp"], x["p"], x["q_t"], x["qp_t"], L, S,
                              dtype=np.float64)
    g = O.step_bwd(P, pat, x["n"], x["p"], x["q_t"], x["qp_t"], L, S, saved, a, w,
                   dtype=np.float64)
    qn32, saved32, a32 = O.step_fwd(P, pat, x["n"], x["qsp"], x["p"], x["q_t"], x["qp_t"], L,
                                    S, dtype=np.float32)
    g32 = O.step_bwd(P, pat, x["n"], x["p"], x["q_t"], x["qp_t"], L, S, saved32, a32, w,
                     dtype=np.float32)
    for par in PARENTS:
        fd = np.zeros(R)
        for i in range(R):
            h = 1e-6 * max(abs(x[par][i]), 1e-3)
            xp = {k: np.array(v, np.float64) for k, v in x.items()}
            xm = {k: np.array(v, np.float64) for k, v in x.items()}
            xp[par][i] += h
            xm[par][i] -= h
            fd[i] = (_step_loss(P, pat, xp, L, S, w) - _step_loss(P, pat, xm, L, S, w)) / (2 * h)
        an = g[GRAD_KEY[par]]
        scale = np.abs(fd).max() + 1e-300
        np.testing.assert_allclose(an, fd, rtol=rel_tol, atol=1e-7 * scale, err_msg=par)
        # the f32 restatement (the reference's arithmetic) within the 1e-3 gate
        np.testing.assert_allclose(g32[GRAD_KEY[par]], an, rtol=1e-3, atol=1e-4 * scale,
                                   err_msg=f"f32 {par}")


def _inputs(d, R):
    return {k: np.array(v, np.float64) for k, v in
            dict(n=d["n"], qsp=d["qsp"], p=d["p"], q_t=d["q_t"], qp_t=d["qp_t"]).items()}


def test_step_gradcheck_sp8_fixture():
    """tests/sp8_gradcheck.rs fixture (4-reach chain, 1 km, ddr_match=false)."""
    R = 4
    pat = O.pattern_from_coo(R, *chain(R))
    _check_step(mock_params(O), pat, _inputs(SP8_INPUTS, R), np.full(R, 1000.0),
                np.full(R, 0.001), 1e-5)


def test_step_gradcheck_cunge_x_fixture():
    """tests/cunge_x.rs:184-199: 5000 m reaches keep X unclamped so B19 is live."""
    R = 4
    pat = O.pattern_from_coo(R, *chain(R))
    P = mock_params(O)
    x = _inputs(SP8_INPUTS, R)
    L, S = np.full(R, 5000.0), np.full(R, 0.001)
    _, saved, _ = O.step_fwd(P, pat, x["n"], x["qsp"], x["p"], x["q_t"], x["qp_t"], L, S,
                             dtype=np.float64)
    xe = saved[O.SAVE["x_eff"]]
    assert ((xe > 0.02) & (xe < 0.48)).all()
    _check_step(P, pat, x, L, S, 1e-5)


def test_step_gradcheck_ddr_match():
    R = 4
    pat = O.pattern_from_coo(R, *chain(R))
    _check_step(mock_params(O, ddr_match=True), pat, _inputs(SP8_INPUTS, R),
                np.full(R, 1000.0), np.full(R, 0.001), 1e-5)


def test_step_gradcheck_positivity_fixture():
    """tests/positivity_clamp.rs:481-485 grad_fixture with the clamp on (B18'/B19')."""
    R = 10
    pat = O.pattern_from_coo(R, *chain(R))
    x = _inputs(dict(n=[0.035] * R, qsp=[0.4] * R, p=[20.0] * R, q_t=POS_QT_GRAD,
                     qp_t=[0.01] * R), R)
    _check_step(stress_params(O, True), pat, x, np.array(POS_LENGTH), np.full(R, 0.001), 1e-4)


@pytest.mark.parametrize("inject_q0", [False, True])
def test_route_adjoint_fd(inject_q0):
    """Whole-window reverse-time adjoint against FP64 central differences on
    seeded reaches, for n, q_spatial, p_spatial, q' and Q0."""
    rng = np.random.default_rng(7)
    R, T = 30, 10
    rows, cols = random_tree(R, 6, rng)
    pat = O.pattern_from_coo(R, rows, cols)
    P = O.make_params()
    base = dict(n=rng.uniform(0.015, 0.25, R), qsp=rng.uniform(0, 1, R), p=np.full(R, 21.0),
                L=rng.uniform(1000, 10000, R),
                S=np.exp(rng.uniform(np.log(1e-4), np.log(5e-2), R)),
                qp=rng.lognormal(np.log(0.5), 1.0, (T, R)),
                q0=rng.lognormal(np.log(5.0), 1.0, R) if inject_q0 else None)
    w = rng.uniform(0.5, 1.5, (T, R))

    def loss(v):
        Q = O.route_fwd(P, pat, v["n"], v["qsp"], v["p"], v["L"], v["S"], v["qp"], q0=v["q0"],
                        dtype=np.float64)
        return float((w * Q).sum())

    g = O.route_adj(P, pat, base["n"], base["qsp"], base["p"], base["L"], base["S"], base["qp"],
                    w, q0=base["q0"], dtype=np.float64)
    checks = [("n", "n", i) for i in range(0, R, 3)] + [("qsp", "q_spatial", i) for i in range(1, R, 4)]
    checks += [("p", "p_spatial", i) for i in range(2, R, 7)]
    checks += [("qp", "q_prime", (t, i)) for t, i in [(0, 3), (2, 5), (4, 11), (7, 20), (8, 29)]]
    if inject_q0:
        checks += [("q0", "q0", i) for i in (0, 4, 9, 17)]
    for key, gkey, idx in checks:
        vp = {k: (None if v is None else np.array(v, np.float64)) for k, v in base.items()}
        vm = {k: (None if v is None else np.array(v, np.float64)) for k, v in base.items()}
        h = 1e-6 * max(abs(vp[key][idx]), 1e-3)
        vp[key][idx] += h
        vm[key][idx] -= h
        fd = (loss(vp) - loss(vm)) / (2 * h)
        an = g[gkey][idx]
        assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-3), (key, idx, an, fd)

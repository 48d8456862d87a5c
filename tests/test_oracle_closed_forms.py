"""Pins the CPU oracle against what the mathematics fixes: closed forms,
identities and mass conservation the reference's tests assert. CPU only."""
import numpy as np

from oracle import oracle as O
from tests.nets import POS_LENGTH, POS_QT, chain, mock_params, random_tree, stress_params


def _pat(R, rows, cols):
    return O.pattern_from_coo(R, rows, cols)


def test_hotstart_chain_is_cumulative_sum():
    """tests/mmc.rs:24-62."""
    pat = _pat(5, *chain(5))
    np.testing.assert_allclose(O.hotstart(O.make_params(), pat, np.full(5, 2.0)),
                               [2, 4, 6, 8, 10], rtol=1e-6)
    pat = _pat(4, *chain(4))
    np.testing.assert_allclose(O.hotstart(O.make_params(), pat, [3.0, 1.0, 2.0, 4.0]),
                               [3, 4, 6, 10], rtol=1e-6)
    pat = _pat(3, *chain(3))
    q0 = O.hotstart(O.make_params(discharge_lb=1e-3), pat, np.full(3, 1e-5))
    assert (q0 >= 1e-3).all()


def _random_step(seed, dtype, **kw):
    rng = np.random.default_rng(seed)
    R = 300
    rows, cols = random_tree(R, 16, rng)
    pat = _pat(R, rows, cols)
    n = rng.uniform(0.015, 0.25, R)
    qsp = rng.uniform(0, 1, R)
    p = np.full(R, 21.0)
    q = rng.lognormal(np.log(20.0), 1.5, R)
    qp = rng.lognormal(np.log(0.5), 1.0, R)
    L = rng.uniform(1000, 10000, R)
    S = np.maximum(np.exp(rng.uniform(np.log(1e-4), np.log(5e-2), R)), 1e-3)
    P = O.make_params(**kw)
    qn, saved, a = O.step_fwd(P, pat, n, qsp, p, q, qp, L, S, dtype=dtype)
    return dict(pat=pat, n=n, qsp=qsp, p=p, q=q, qp=qp, L=L, S=S, P=P, qn=qn, saved=saved, a=a)


def test_coefficients_partition_identity():
    """c1 + c2 + c3 = 1 for any (K, X) (tests/cunge_x.rs:128-140)."""
    d = _random_step(1, np.float64)
    s = d["saved"]
    np.testing.assert_allclose(s[O.SAVE["c1"]] + s[O.SAVE["c2"]] + s[O.SAVE["c3"]], 1.0, atol=1e-12)
    f = _random_step(1, np.float32)["saved"]
    np.testing.assert_allclose(f[O.SAVE["c1"]] + f[O.SAVE["c2"]] + f[O.SAVE["c3"]], 1.0, atol=2e-5)


def _manning_q(b, z, y):
    a = (b + z * y) * y
    p = b + 2 * y * np.sqrt(1 + z * z)
    return a ** (5 / 3) * p ** (-2 / 3)


def test_celerity_is_trapezoid_dq_da():
    """S17 beta: celerity / v = dQ/dA / (Q/A) of Manning flow in the trapezoid the
    step builds (tests/celerity_beta.rs:57-105), on reaches where no geometry
    clamp binds; beta lies in (1, 5/3) (138-156)."""
    d = _random_step(2, np.float64)
    s = d["saved"]
    y, z, b = s[O.SAVE["depth"]], s[O.SAVE["side_slope"]], s[O.SAVE["bottom_width"]]
    free = ((s[O.SAVE["side_slope_raw"]] > 0.5) & (s[O.SAVE["side_slope_raw"]] < 50)
            & (s[O.SAVE["bw_raw"]] > 0.01) & (s[O.SAVE["v_un"]] > 0.01) & (s[O.SAVE["v_un"]] < 15))
    assert free.sum() > 50
    beta = s[O.SAVE["celerity"]] / s[O.SAVE["v_cl"]]
    h = 1e-6 * y
    dq = _manning_q(b, z, y + h) - _manning_q(b, z, y - h)
    da = (b + z * (y + h)) * (y + h) - (b + z * (y - h)) * (y - h)
    q = _manning_q(b, z, y)
    a = (b + z * y) * y
    beta_fd = (dq / da) / (q / a)
    np.testing.assert_allclose(beta[free], beta_fd[free], rtol=1e-6)
    assert (beta > 1).all() and (beta < 5 / 3).all()


def test_cunge_x_matches_physical_diffusion():
    """S19: on unclamped reaches D_num = c L (0.5 - X) equals D_phys = Q/(2 T S)
    (tests/cunge_x.rs:35-78)."""
    d = _random_step(3, np.float64)
    s = d["saved"]
    x = s[O.SAVE["x_eff"]]
    inner = (x > 0) & (x < 0.5)
    assert inner.sum() > 20
    dnum = s[O.SAVE["celerity"]] * d["L"] * (0.5 - x)
    dphys = d["q"] / (2 * s[O.SAVE["top_width"]] * d["S"])
    np.testing.assert_allclose(dnum[inner], dphys[inner], rtol=1e-8)


def test_ddr_match_branch_constants():
    """ddr_match: celerity = 5/3 v and X = x_storage (mmc_op.rs:1155-1162,1196-1200)."""
    d = _random_step(4, np.float64, ddr_match=True)
    s = d["saved"]
    np.testing.assert_allclose(s[O.SAVE["celerity"]], s[O.SAVE["v_cl"]] * 5 / 3, rtol=1e-14)
    np.testing.assert_allclose(s[O.SAVE["x_eff"]], 0.3)


def test_solve_matches_dense_solve():
    """S27 against a dense solve of (I - c1 N) x = b (brute force, f64)."""
    d = _random_step(5, np.float64)
    pat, s = d["pat"], d["saved"]
    R = pat.n
    A = np.eye(R)
    rows = pat.row_for_nnz[pat.diag_mask == 0]
    cols = pat.col[pat.diag_mask == 0]
    A[rows, cols] -= s[O.SAVE["c1"]][rows]
    x = np.linalg.solve(A, s[O.SAVE["b_rhs"]])
    np.testing.assert_allclose(s[O.SAVE["x_sol"]], x, rtol=1e-10, atol=1e-10)
    i_t = np.zeros(R)
    np.add.at(i_t, rows, d["q"][cols])
    np.testing.assert_allclose(s[O.SAVE["i_t"]], i_t, rtol=1e-12)


def test_positivity_clamp_keeps_courant_below_two():
    """S18' floors K at dt(1+d)/2 (tests/positivity_clamp.rs:398-430)."""
    R = 10
    pat = _pat(R, *chain(R))
    args = (np.full(R, 0.035), np.full(R, 0.4), np.full(R, 20.0), POS_QT, np.full(R, 0.01),
            POS_LENGTH, np.full(R, 0.001))
    _, on, _ = O.step_fwd(stress_params(O, True), pat, *args, dtype=np.float64)
    k = on[O.SAVE["k_musk"]]
    assert (k >= 3600 * 1.01 / 2 * (1 - 1e-12)).all()
    assert (3600 / k <= 2 / 1.01 * (1 + 1e-12)).all()


def test_steady_state_mass_conservation():
    """Two headwaters into a gauge reach: at steady state the gauge carries the
    sum of q' over the basin (tests/gauge_mass_conservation.rs:192-221)."""
    R, T = 3, 64
    pat = _pat(R, np.array([2, 2], np.int32), np.array([0, 1], np.int32))
    P = mock_params(O)
    qp = np.full((T, R), 10.0)
    Q = O.route_fwd(P, pat, np.full(R, 0.055), np.full(R, 0.5), np.full(R, 1.0),
                    np.full(R, 1000.0), np.full(R, 0.001), qp, dtype=np.float64)
    assert abs(Q[-1, 2] - 30.0) / 30.0 < 1e-2
    assert abs(Q[-1, 0] - 10.0) / 10.0 < 1e-2

"""Small network fixtures shared by the tests (restated from the reference's
test files, with the file:line each one follows)."""
import numpy as np


def chain(n):
    """Linear chain N[i+1, i] = 1 (tests/common.rs:45-52)."""
    return np.arange(1, n, dtype=np.int32), np.arange(0, n - 1, dtype=np.int32)


def random_tree(R, window, rng):
    """Reach i < R-1 drains to a reach drawn from (i, min(i+window, R-1)]."""
    i = np.arange(R - 1)
    hi = np.minimum(i + window, R - 1)
    down = i + 1 + np.floor(rng.random(R - 1) * (hi - i)).astype(np.int64)
    return down.astype(np.int32), i.astype(np.int32)


# tests/positivity_clamp.rs:111-125 (fixture) and 481-485 (grad_fixture)
POS_LENGTH = [400.0, 800.0, 1500.0, 2500.0, 3600.0, 5000.0, 8000.0, 20000.0, 60000.0, 150000.0]
POS_QT = [6.0, 8.0, 15.0, 50.0, 120.0, 200.0, 100.0, 100.0, 100.0, 100.0]
POS_QT_GRAD = [6.0, 8.0, 15.0, 50.0, 120.0, 200.0, 300.0, 40.0, 400.0, 30.0]
# tests/positivity_clamp.rs:59-70
GOLDEN_X_SOL_BITS = [0xC0645CCA, 0xBF9F4594, 0x3F511646, 0xBF7F58E2, 0x4143D4D4,
                     0x42857D42, 0x433334A2, 0x428D80F3, 0x42F2EE24, 0x42A20A08]


def stress_params(O, enforce_positivity):
    """stress_cfg (tests/positivity_clamp.rs:127-144)."""
    return O.make_params(velocity_lb=0.01, depth_lb=0.001, discharge_lb=1e-4,
                         bottom_width_lb=0.01, slope_lb=1e-4,
                         enforce_positivity=enforce_positivity)


def mock_params(O, ddr_match=False):
    """mock_cfg (tests/sp8_gradcheck.rs:45-58, tests/cunge_x.rs:200-215)."""
    return O.make_params(velocity_lb=0.1, depth_lb=0.01, discharge_lb=0.001,
                         bottom_width_lb=0.1, slope_lb=0.001, ddr_match=ddr_match)


# tests/sp8_gradcheck.rs:68-75
SP8_INPUTS = dict(n=[0.035] * 4, qsp=[0.4] * 4, p=[20.0] * 4,
                  q_t=[100.0, 120.0, 140.0, 160.0], qp_t=[10.0, 12.0, 14.0, 16.0])

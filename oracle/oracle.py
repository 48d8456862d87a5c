"""ctypes bindings of the CPU oracle — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker. It wraps oracle/liboracle.so (built by
oracle/Makefile), a plain-C restatement of the reference ddrs CPU path for the
Muskingum-Cunge timestep (float and double variants).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NSAVE = 23
SAVE_NAMES = [
    "q_eps", "ratio", "denominator", "depth", "top_width", "side_slope_raw",
    "side_slope", "bw_raw", "bottom_width", "hyd_radius", "v_un", "v_cl",
    "celerity", "k_musk", "x_eff", "denom", "c1", "c2", "c3", "c4", "i_t",
    "b_rhs", "x_sol",
]
SAVE = {name: k for k, name in enumerate(SAVE_NAMES)}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("discharge_lb", ctypes.c_double),
        ("depth_lb", ctypes.c_double),
        ("velocity_lb", ctypes.c_double),
        ("bottom_width_lb", ctypes.c_double),
        ("slope_lb", ctypes.c_double),
        ("ddr_match", ctypes.c_int),
        ("enforce_positivity", ctypes.c_int),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS: dict[str, ctypes.CDLL] = {}


def lib() -> ctypes.CDLL:
    """Load oracle/liboracle.so, building it first if needed."""
    name = "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        _LIBS[name] = ctypes.CDLL(path)
    return _LIBS[name]


def set_threads(n: int) -> None:
    """Threads of the per-reach loops of the oracle's step functions (1: serial,
    the default). Results do not depend on it."""
    lib().orc_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def make_params(dt=3600.0, discharge_lb=1e-4, depth_lb=0.01, velocity_lb=0.01,
                bottom_width_lb=0.01, slope_lb=1e-3, ddr_match=False,
                enforce_positivity=False) -> OrcParams:
    """Defaults are the reference's (config.rs AttributeMinimums 436-446,
    Params::default 652-672)."""
    return OrcParams(dt, discharge_lb, depth_lb, velocity_lb, bottom_width_lb, slope_lb,
                     int(ddr_match), int(enforce_positivity))


@dataclass
class Pattern:
    n: int
    crow: np.ndarray
    col: np.ndarray
    row_for_nnz: np.ndarray
    adj_values: np.ndarray
    diag_mask: np.ndarray
    trans_crow: np.ndarray
    trans_col: np.ndarray
    trans_to_orig: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])


def pattern_from_coo(n: int, rows, cols, vals=None) -> Pattern:
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ones(rows.shape[0], np.float32) if vals is None else np.ascontiguousarray(vals, np.float32)
    nnz = rows.shape[0] + n
    out = dict(
        crow=np.zeros(n + 1, np.int32), col=np.zeros(nnz, np.int32),
        row_for_nnz=np.zeros(nnz, np.int32), adj_values=np.zeros(nnz, np.float32),
        diag_mask=np.zeros(nnz, np.float32), trans_crow=np.zeros(n + 1, np.int32),
        trans_col=np.zeros(nnz, np.int32), trans_to_orig=np.zeros(nnz, np.int32),
    )
    rc = lib().orc_pattern_from_coo(
        ctypes.c_int(n), ctypes.c_int(rows.shape[0]), _p(rows), _p(cols), _p(vals),
        _p(out["crow"]), _p(out["col"]), _p(out["row_for_nnz"]), _p(out["adj_values"]),
        _p(out["diag_mask"]), _p(out["trans_crow"]), _p(out["trans_col"]), _p(out["trans_to_orig"]))
    if rc != 0:
        raise ValueError("adjacency is not strictly lower triangular")
    return Pattern(n=n, **out)


# dtype of the float variant with the device's correctly rounded powers and
# logarithms (orc_*_c) in place of glibc's powf / logf
F32_DEVICE_MATH = "float32-device-math"


def _dt(dtype):
    if isinstance(dtype, str) and dtype == F32_DEVICE_MATH:
        return "c", np.float32
    return ("f", np.float32) if np.dtype(dtype) == np.float32 else ("d", np.float64)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def step_fwd(P: OrcParams, pat: Pattern, n, qsp, p, q_t, qp_t, length, slope, xst=None,
             dtype=np.float32):
    suf, dt = _dt(dtype)
    R = pat.n
    xst = np.full(R, 0.3) if xst is None else xst
    args = [_c(v, dt) for v in (n, qsp, p, q_t, qp_t, length, slope, xst)]
    adj, dmask = _c(pat.adj_values, dt), _c(pat.diag_mask, dt)
    saved = np.zeros((NSAVE, R), dt)
    a = np.zeros(pat.nnz, dt)
    q_next = np.zeros(R, dt)
    getattr(lib(), f"orc_step_fwd_{suf}")(
        ctypes.byref(P), R, pat.nnz, _p(pat.crow), _p(pat.col), _p(pat.row_for_nnz), _p(adj),
        _p(dmask), *[_p(x) for x in args], _p(saved), _p(a), _p(q_next))
    return q_next, saved, a


def step_bwd(P: OrcParams, pat: Pattern, n, p, q_t, qp_t, length, slope, saved, a, gy,
             dtype=np.float32):
    suf, dt = _dt(dtype)
    R = pat.n
    adj = _c(pat.adj_values, dt)
    ins = [_c(v, dt) for v in (n, p, q_t, qp_t, length, slope)]
    outs = [np.zeros(R, dt) for _ in range(5)]
    getattr(lib(), f"orc_step_bwd_{suf}")(
        ctypes.byref(P), R, pat.nnz, _p(pat.col), _p(pat.row_for_nnz), _p(adj),
        _p(pat.trans_crow), _p(pat.trans_col), _p(pat.trans_to_orig), *[_p(x) for x in ins],
        _p(_c(saved, dt)), _p(_c(a, dt)), _p(_c(gy, dt)), *[_p(o) for o in outs])
    return dict(zip(["n", "q_spatial", "p_spatial", "q_t", "q_prime_t"], outs))


def forward_sub_lower(pat: Pattern, a, b, dtype=np.float32):
    suf, dt = _dt(dtype)
    x = np.zeros(pat.n, dt)
    getattr(lib(), f"orc_forward_sub_lower_{suf}")(pat.n, _p(pat.crow), _p(pat.col),
                                                   _p(_c(a, dt)), _p(_c(b, dt)), _p(x))
    return x


def back_sub_upper_transposed(pat: Pattern, a, b, dtype=np.float32):
    suf, dt = _dt(dtype)
    y = np.zeros(pat.n, dt)
    getattr(lib(), f"orc_back_sub_upper_transposed_{suf}")(
        pat.n, _p(pat.trans_crow), _p(pat.trans_col), _p(pat.trans_to_orig), _p(_c(a, dt)),
        _p(_c(b, dt)), _p(y))
    return y


def hotstart(P: OrcParams, pat: Pattern, qp0, dtype=np.float32):
    suf, dt = _dt(dtype)
    q0 = np.zeros(pat.n, dt)
    getattr(lib(), f"orc_hotstart_{suf}")(
        ctypes.byref(P), pat.n, pat.nnz, _p(pat.crow), _p(pat.col), _p(pat.row_for_nnz),
        _p(_c(pat.adj_values, dt)), _p(_c(pat.diag_mask, dt)), _p(_c(qp0, dt)), _p(q0))
    return q0


def denormalize(v, lo, hi, log_space, dtype=np.float32):
    suf, dt = _dt(dtype)
    v = _c(v, dt)
    out = np.zeros_like(v)
    getattr(lib(), f"orc_denormalize_{suf}")(v.shape[0], _p(v), ctypes.c_double(lo),
                                             ctypes.c_double(hi), int(log_space), _p(out))
    return out


def route_fwd(P: OrcParams, pat: Pattern, n, qsp, p, length, slope_raw, qprime, q0=None,
              xst=None, dtype=np.float32, want_x=False):
    """Q[T][R] (and pre-clamp X[T][R]) of MuskingumCunge::forward."""
    suf, dt = _dt(dtype)
    R = pat.n
    qprime = _c(qprime, dt)
    T = qprime.shape[0]
    xst = np.full(R, 0.3) if xst is None else xst
    ins = [_c(v, dt) for v in (n, qsp, p, length, slope_raw, xst)]
    Q = np.zeros((T, R), dt)
    X = np.zeros((T, R), dt) if want_x else None
    getattr(lib(), f"orc_route_fwd_{suf}")(
        ctypes.byref(P), R, pat.nnz, T, _p(pat.crow), _p(pat.col), _p(pat.row_for_nnz),
        _p(_c(pat.adj_values, dt)), _p(_c(pat.diag_mask, dt)), *[_p(x) for x in ins],
        _p(qprime), _p(_c(q0, dt) if q0 is not None else None), _p(Q), _p(X))
    return (Q, X) if want_x else Q


def route_adj(P: OrcParams, pat: Pattern, n, qsp, p, length, slope_raw, qprime, gQ, q0=None,
              xst=None, dtype=np.float32, want_q=False, want_x=False):
    """Gradients of the loss whose dL/dQ is gQ [T][R] (orc_route_adj); with
    want_q / want_x also the forward's Q and pre-clamp X [T][R] ("Q", "X")."""
    suf, dt = _dt(dtype)
    R = pat.n
    qprime = _c(qprime, dt)
    T = qprime.shape[0]
    xst = np.full(R, 0.3) if xst is None else xst
    ins = [_c(v, dt) for v in (n, qsp, p, length, slope_raw, xst)]
    gn, gqsp, gp, gq0 = (np.zeros(R, dt) for _ in range(4))
    gqp = np.zeros((T, R), dt)
    Qo = np.zeros((T, R), dt) if want_q else None
    Xo = np.zeros((T, R), dt) if want_x else None
    rc = getattr(lib(), f"orc_route_adj_{suf}")(
        ctypes.byref(P), R, pat.nnz, T, _p(pat.crow), _p(pat.col), _p(pat.row_for_nnz),
        _p(_c(pat.adj_values, dt)), _p(_c(pat.diag_mask, dt)), _p(pat.trans_crow),
        _p(pat.trans_col), _p(pat.trans_to_orig), *[_p(x) for x in ins], _p(qprime),
        _p(_c(q0, dt) if q0 is not None else None), _p(_c(gQ, dt)), _p(gn), _p(gqsp), _p(gp),
        _p(gqp), _p(gq0), _p(Qo), _p(Xo))
    if rc != 0:
        raise MemoryError("oracle adjoint allocation failed")
    out = dict(n=gn, q_spatial=gqsp, p_spatial=gp, q_prime=gqp, q0=gq0)
    if want_q:
        out["Q"] = Qo
    if want_x:
        out["X"] = Xo
    return out


def downstream_path(pat: Pattern, r: int) -> np.ndarray:
    """Reach r and every reach downstream of it, in flow order (each reach's
    downstream reach is the row of its single off-diagonal entry in the
    transposed pattern; a reach with none is an outlet)."""
    path = [int(r)]
    while True:
        i = path[-1]
        down = [int(j) for j in pat.trans_col[pat.trans_crow[i]:pat.trans_crow[i + 1]] if j != i]
        if not down:
            return np.asarray(path, np.int32)
        if len(down) > 1:
            raise ValueError(f"reach {i} drains to {len(down)} reaches; not a tree")
        path.append(down[0])


def route_path(P: OrcParams, pat: Pattern, n, qsp, p, length, slope_raw, qprime, Qb, Xb, path,
               xst=None, dtype=np.float64):
    """Discharge [T][len(path)] of the reaches of a downstream path after the
    parameters of path[0] changed (orc_route_path): Qb / Xb are the unchanged
    run's discharge and pre-clamp solve, every other reach keeps its values."""
    suf, dt = _dt(dtype)
    R = pat.n
    qprime = _c(qprime, dt)
    T = qprime.shape[0]
    xst = np.full(R, 0.3) if xst is None else xst
    ins = [_c(v, dt) for v in (n, qsp, p, length, slope_raw, xst)]
    path = np.ascontiguousarray(path, np.int32)
    Qp = np.zeros((T, path.shape[0]), dt)
    rc = getattr(lib(), f"orc_route_path_{suf}")(
        ctypes.byref(P), R, T, _p(pat.crow), _p(pat.col), _p(_c(pat.adj_values, dt)),
        _p(_c(pat.diag_mask, dt)), *[_p(x) for x in ins], _p(qprime), _p(_c(Qb, dt)),
        _p(_c(Xb, dt)), path.shape[0], _p(path), _p(Qp))
    if rc != 0:
        raise ValueError("route_path: the path is not a downstream chain")
    return Qp

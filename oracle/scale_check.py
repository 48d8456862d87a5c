"""Workload-size parity against the oracle — test infrastructure only.

Used by tests/ and by bench.py's check leg, only as the checker: every function
takes device results already copied to numpy and runs the oracle (oracle.py)
on the same inputs. Nothing here touches the GPU.

route_pair     the oracle's forward + adjoint of one network in float (the
               reference's arithmetic), double, and float with the device's
               powers, in three threads (ctypes releases the GIL), with the
               same dL/dQ (route_fwd_set: the forward alone)
discharge_errors / grad_errors
               the parity figures: max relative discharge error against the
               f32 oracle (everywhere, and where the f32 oracle is itself
               within 1e-4 of FP64) and against FP64, and per gradient vector the
               norm-wise error (max|g - g_ref| / max|g_ref|) plus the
               element-wise relative error on the well-conditioned entries
               (|g_ref| >= 1e-2 max|g_ref|)
fd_check       FP64 central finite differences of a linear loss
               L = sum_t sum_g w[t, g] Q[t, g] on seeded reaches, each
               perturbed run re-routing only the reach's downstream path
               (oracle.route_path: exact, the other reaches do not move)
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import oracle as O

GRAD_KEYS = ("n", "q_spatial")


def threads() -> int:
    """Worker threads for the oracle (at most 16: the CPU share of one GPU)."""
    return max(1, min(16, os.cpu_count() or 1))


def route_pair(P, pat, n, qsp, p, length, slope, qprime, gQ, q0, want_x64=False, pool=None):
    """{"f32", "f64", "dev"}: oracle.route_adj in float (glibc's powers, the
    reference's arithmetic), in double, and in float with the device's
    correctly rounded powers, in three threads, each with the forward's Q
    (and, with want_x64, the double pre-clamp X for fd_check)."""
    q64 = np.ascontiguousarray(qprime, np.float64)
    own = pool is None
    pool = pool or ThreadPoolExecutor(3)
    try:
        f32 = pool.submit(O.route_adj, P, pat, n, qsp, p, length, slope, qprime, gQ, q0=q0,
                          want_q=True)
        dev = pool.submit(O.route_adj, P, pat, n, qsp, p, length, slope, qprime, gQ, q0=q0,
                          want_q=True, dtype=O.F32_DEVICE_MATH)
        f64 = pool.submit(O.route_adj, P, pat, n, qsp, p, length, slope, q64,
                          np.asarray(gQ, np.float64), q0=np.asarray(q0, np.float64),
                          dtype=np.float64, want_q=True, want_x=want_x64)
        return {"f32": f32.result(), "f64": f64.result(), "dev": dev.result()}
    finally:
        if own:
            pool.shutdown()


def route_fwd_set(P, pat, n, qsp, p, length, slope, qprime, q0, pool=None):
    """route_pair's three variants, forward only: {variant: {"Q": [T][R]}}."""
    own = pool is None
    pool = pool or ThreadPoolExecutor(3)
    try:
        f = lambda dt, qp, q: pool.submit(O.route_fwd, P, pat, n, qsp, p, length, slope, qp,
                                          q0=q, dtype=dt)
        futs = {"f32": f(np.float32, qprime, q0), "dev": f(O.F32_DEVICE_MATH, qprime, q0),
                "f64": f(np.float64, np.ascontiguousarray(qprime, np.float64),
                         np.asarray(q0, np.float64))}
        return {k: {"Q": v.result()} for k, v in futs.items()}
    finally:
        if own:
            pool.shutdown()


def discharge_errors(got, r: dict, cols=None) -> dict:
    """Relative discharge errors of the device against the three oracles of
    route_pair (optionally at columns `cols` of their [T][R] outputs), and the
    spread of the f32 oracle against the device-math one: the part of the
    reference's f32 result that libm's rounding of a small fraction of powers
    decides. Where the recurrence amplifies single roundings (with Cunge X most
    reaches have c1 < 0), two f32 evaluations that round one power differently
    part there by more than 1e-4 while both stay as far from FP64."""
    take = (lambda a: a) if cols is None else (lambda a: a[:, cols])
    got = np.asarray(got, np.float64)
    ref32, ref64, dev = (np.asarray(take(r[k]["Q"]), np.float64) for k in ("f32", "f64", "dev"))
    r32 = np.abs(got - ref32) / np.abs(ref32)
    rdev = np.abs(got - dev) / np.abs(dev)
    spread = np.abs(ref32 - dev) / np.abs(dev)
    over = r32 > 1e-4
    return {"max_rel_err_vs_f32_oracle": float(r32.max()),
            "frac_within_1e-4": float(1.0 - over.mean()),
            "count_over_1e-4": int(over.sum()),
            "bit_identical_to_device_math_oracle": float((got == dev).mean()),
            "max_rel_err_vs_device_math_oracle": float(rdev.max()),
            "f32_oracle_vs_device_math_oracle_max_rel": float(spread.max()),
            "f32_oracle_vs_device_math_oracle_count_over_1e-4": int((spread > 1e-4).sum()),
            "max_rel_err_vs_fp64": float((np.abs(got - ref64) / np.abs(ref64)).max()),
            "f32_oracle_max_rel_err_vs_fp64": float((np.abs(ref32 - ref64) / np.abs(ref64)).max()),
            "values": int(got.size)}


def grad_errors(got, r: dict, key: str, idx=None, well=1e-2) -> dict:
    """Gradient figures of the device's `got` against route_pair's r[.][key]
    (optionally at entries idx)."""
    take = (lambda a: a) if idx is None else (lambda a: a[idx])
    got = np.asarray(got, np.float64)
    ref32, ref64, dev = (np.asarray(take(r[k][key]), np.float64) for k in ("f32", "f64", "dev"))
    s32 = max(float(np.abs(ref32).max()), 1e-300)
    s64 = max(float(np.abs(ref64).max()), 1e-300)
    wc = np.abs(ref64) >= well * s64
    rel = lambda a, b: np.abs(a - b)[wc] / np.abs(b)[wc]
    rel32, rel64, orc64 = rel(got, ref32), rel(got, ref64), rel(ref32, ref64)
    m = lambda a: float(a.max()) if a.size else 0.0
    return {"norm_err_vs_f32_oracle": float(np.abs(got - ref32).max()) / s32,
            "norm_err_vs_device_math_oracle": float(np.abs(got - dev).max()) / s32,
            "f32_oracle_vs_device_math_oracle_norm": float(np.abs(ref32 - dev).max()) / s32,
            "norm_err_vs_fp64": float(np.abs(got - ref64).max()) / s64,
            "f32_oracle_norm_err_vs_fp64": float(np.abs(ref32 - ref64).max()) / s64,
            "well_conditioned_entries": int(wc.sum()), "entries": int(got.size),
            "elem_max_rel_err_vs_f32_oracle": m(rel32),
            "elem_frac_within_1e-3_vs_f32_oracle": float((rel32 <= 1e-3).mean()) if rel32.size else 1.0,
            "elem_max_rel_err_vs_device_math_oracle": m(rel(got, dev)),
            "elem_max_rel_err_vs_fp64": m(rel64),
            "f32_oracle_elem_max_rel_err_vs_fp64": m(orc64)}


def fd_check(P, pat, params, length, slope, qprime, Q64, X64, w, gauges, reaches,
             keys=GRAD_KEYS, steps=(1e-3, 3e-4, 1e-4, 3e-5, 1e-5), pool=None) -> dict:
    """FP64 central differences of L = sum_t sum_k w[t, k] Q[t, gauges[k]] with
    respect to params[key][r] for every r in reaches and key in keys.
    params: {"n", "q_spatial", "p_spatial"} float64 arrays of the base point;
    Q64 / X64: the double oracle's run at that point.

    The step is chosen per reach. Small steps drown in double rounding, which
    the recurrence amplifies by many orders of magnitude on deep networks; large
    ones can straddle the kink of a clamp (a reach-step crossing the discharge,
    depth, velocity or Cunge-X bound). So the difference is taken at every
    relative step in `steps`, and the adjacent pair that agrees best gives the
    value (their mean) and the consistency (their relative difference).
    Returns {key: {"fd": values in the order of reaches, "consistency": ...,
    "step": the smaller step of the chosen pair}}."""
    q64 = np.ascontiguousarray(qprime, np.float64)
    base = {k: np.ascontiguousarray(v, np.float64) for k, v in params.items()}
    slot = {int(g): i for i, g in enumerate(np.asarray(gauges))}
    w = np.asarray(w, np.float64)

    def central(r, key, path, on, wk, rel):
        h = rel * abs(base[key][r])
        outs = []
        for s in (1.0, -1.0):
            v = dict(base)
            v[key] = base[key].copy()
            v[key][r] += s * h
            Qp = O.route_path(P, pat, v["n"], v["q_spatial"], v["p_spatial"], length, slope, q64,
                              Q64, X64, path)
            outs.append(Qp[:, on])
        return float((wk * (outs[0] - outs[1])).sum()) / (2.0 * h)

    def one(r, key):
        path = O.downstream_path(pat, r)
        on = [k for k, g in enumerate(path) if int(g) in slot]
        if not on:
            return 0.0, 0.0, steps[-1]
        wk = w[:, [slot[int(path[k])] for k in on]]
        f = [central(r, key, path, on, wk, rel) for rel in steps]
        best = None
        for k in range(len(f) - 1):
            c = abs(f[k] - f[k + 1]) / max(abs(f[k + 1]), 1e-300)
            if best is None or c < best[0]:
                best = (c, 0.5 * (f[k] + f[k + 1]), steps[k + 1])
        return best[1], best[0], best[2]

    own = pool is None
    pool = pool or ThreadPoolExecutor(threads())
    try:
        futs = {key: [pool.submit(one, int(r), key) for r in reaches] for key in keys}
        out = {}
        for key, fs in futs.items():
            res = [f.result() for f in fs]
            out[key] = {"fd": np.array([a for a, _, _ in res]),
                        "consistency": np.array([b for _, b, _ in res]),
                        "step": np.array([c for _, _, c in res])}
        return out
    finally:
        if own:
            pool.shutdown()


def fd_errors(fd: dict, reaches, got: dict, r: dict, floor=1e-6, agree=1e-3) -> dict:
    """Errors against the finite differences, per gradient. A reach counts as
    resolved when its two chosen steps agree within `agree` and its difference
    is at least floor x the largest over the reaches (below that the gradient
    vanishes through thousands of damping links to the nearest gauge, under
    what double differences resolve). Errors are relative to the difference on
    resolved reaches and to the largest difference elsewhere."""
    out = {}
    for key, d in fd.items():
        f = d["fd"]
        top = max(float(np.abs(f).max()), 1e-300)
        res = (np.abs(f) >= floor * top) & (d["consistency"] <= agree)
        den = np.where(res, np.abs(f), top)
        e = lambda g: np.abs(np.asarray(g, np.float64)[reaches] - f) / den
        eg, e32, e64 = e(got[key]), e(r["f32"][key]), e(r["f64"][key])
        pick = lambda a: float(a[res].max()) if res.any() else 0.0
        out[key] = {"reaches": int(len(f)), "resolved": int(res.sum()),
                    "vanishing": int((np.abs(f) < floor * top).sum()),
                    "max_step_consistency_resolved": pick(d["consistency"]),
                    "max_rel_err_gpu": pick(eg),
                    "max_rel_err_f32_oracle": pick(e32),
                    "max_rel_err_fp64_adjoint": pick(e64),
                    "unresolved_max_err_over_max_fd": float(np.r_[eg[~res], 0.0].max()),
                    "resolved_mask": [bool(x) for x in res],
                    "gpu_rel_err_per_reach": [float(x) for x in eg],
                    "f32_oracle_rel_err_per_reach": [float(x) for x in e32],
                    "fp64_adjoint_rel_err_per_reach": [float(x) for x in e64]}
    return out

"""Networks and workloads: seeded synthetic generators and the reference's
index construction (topological order, batch compression), all native."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N
from .sparse import SparseAdjacency, _ptr, _stream


def gen_tree(n: int, window: int = 256, seed: int = 1001):
    """Reach i < n-1 drains to a reach drawn uniformly from (i, min(i+window, n-1)]."""
    rows = np.zeros(max(n - 1, 0), np.int32)
    cols = np.zeros(max(n - 1, 0), np.int32)
    N.check(N.lib.ddrs_gen_tree(n, window, seed, _ptr(rows), _ptr(cols)), "ddrs_gen_tree")
    return rows, cols


def gen_forest(sizes, window: int = 256, seed: int = 1001):
    """Independent basins side by side: basin b is gen_tree(sizes[b], window,
    seed + b) on positions [offset_b, offset_b + sizes[b]). Returns (rows, cols,
    outlets), outlets[b] the last position of basin b. Still lower-triangular
    (block diagonal)."""
    rows, cols, outlets = [], [], []
    off = 0
    for b, n in enumerate(int(x) for x in sizes):
        r, c = gen_tree(n, window, seed + b)
        rows.append(r + off)
        cols.append(c + off)
        off += n
        outlets.append(off - 1)
    cat = lambda a: np.concatenate(a).astype(np.int32) if a else np.zeros(0, np.int32)
    return cat(rows), cat(cols), np.asarray(outlets, np.int32)


def gen_attributes(n: int, seed: int = 2001):
    out = [np.zeros(n, np.float32) for _ in range(4)]
    N.check(N.lib.ddrs_gen_attributes(n, seed, *[_ptr(o) for o in out]), "ddrs_gen_attributes")
    length, slope, nman, qsp = out
    return dict(length=length, slope=slope, n=nman, q_spatial=qsp)


def gen_forcing(n: int, T: int, seed: int = 3001, storms: bool = False) -> np.ndarray:
    q = np.zeros((T, n), np.float32)
    N.check(N.lib.ddrs_gen_forcing(n, T, seed, int(storms), _ptr(q)), "ddrs_gen_forcing")
    return q


def gen_forcing_device(n: int, T: int, seed: int = 3001, storms: bool = False,
                       device="cuda") -> torch.Tensor:
    q = torch.empty((T, n), dtype=torch.float32, device=device)
    N.check(N.lib.ddrs_gen_forcing_device(n, T, seed, int(storms), _ptr(q), _stream()),
            "ddrs_gen_forcing_device")
    return q


def basin_tree(n: int, window: int = 256, seed: int = 1001) -> SparseAdjacency:
    """A synthetic single-outlet river tree with generated length and slope."""
    rows, cols = gen_tree(n, window, seed)
    attrs = gen_attributes(n, seed + 1000)
    return SparseAdjacency.from_coo(n, rows, cols, None, attrs["length"], attrs["slope"])


def build_topology(records: np.ndarray):
    """records [n, 5] int64 (comid, up1..up4) -> (order, rows, cols, dropped),
    adjacency/build.rs:170-272."""
    rec = np.ascontiguousarray(records, np.int64).reshape(-1, 5)
    n = rec.shape[0]
    cap = 5 * n + 1
    order = np.zeros(cap, np.int64)
    rows = np.zeros(cap, np.int32)
    cols = np.zeros(cap, np.int32)
    dropped = np.zeros(cap, np.int64)
    cnt = [ctypes.c_int32() for _ in range(3)]
    N.check(N.lib.ddrs_build_topology(n, _ptr(rec), _ptr(order), cap, ctypes.byref(cnt[0]),
                                      _ptr(rows), _ptr(cols), cap, ctypes.byref(cnt[1]),
                                      _ptr(dropped), cap, ctypes.byref(cnt[2])),
            "ddrs_build_topology")
    no, ne, nd = (c.value for c in cnt)
    return order[:no], rows[:ne], cols[:ne], dropped[:nd]


def compress(rows, cols, gauges):
    """Unioned CONUS COO + gauge positions -> compressed positions
    (data/collate.rs:197-349). Returns (active, rows, cols, gauge_compressed)."""
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    gauges = np.ascontiguousarray(gauges, np.int32)
    cap = 2 * rows.shape[0] + gauges.shape[0] + 1
    active = np.zeros(cap, np.int32)
    orow = np.zeros(max(rows.shape[0], 1), np.int32)
    ocol = np.zeros(max(rows.shape[0], 1), np.int32)
    og = np.zeros(max(gauges.shape[0], 1), np.int32)
    na, ne = ctypes.c_int32(), ctypes.c_int32()
    N.check(N.lib.ddrs_compress(rows.shape[0], _ptr(rows), _ptr(cols), gauges.shape[0],
                                _ptr(gauges), _ptr(active), cap, ctypes.byref(na), _ptr(orow),
                                _ptr(ocol), ctypes.byref(ne), _ptr(og)), "ddrs_compress")
    return active[:na.value], orow[:ne.value], ocol[:ne.value], og[:gauges.shape[0]]


@dataclass
class Compressed:
    """collate::compress's outputs (data/collate.rs:52-104): the sorted active
    set (CONUS sub-reach rows, `conus_positions`), the compressed
    lower-triangular COO, each gauge's compressed position, outflow_idx (gauge
    g sums the discharge of reaches outflow[outflow_ptr[g]:outflow_ptr[g+1]])
    and the compressed reach-subdivision map (None without a CONUS map)."""

    active: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    gauge_compressed: np.ndarray
    outflow_ptr: np.ndarray
    outflow: np.ndarray
    parent_offset: Optional[np.ndarray]

    def outflow_idx(self):
        """outflow_idx as a list of arrays, one per gauge."""
        return [self.outflow[self.outflow_ptr[g]:self.outflow_ptr[g + 1]]
                for g in range(self.gauge_compressed.size)]

    def scatter_map(self):
        """(reaches, flat, group): the distinct reaches a batch must read out
        (sorted), and for every outflow entry its index into them and its
        gauge, the inputs of scatter_add_by_group (training/forward.rs:18-34)."""
        reaches = np.unique(self.outflow).astype(np.int32)
        flat = np.searchsorted(reaches, self.outflow).astype(np.int32)
        group = np.repeat(np.arange(self.gauge_compressed.size, dtype=np.int32),
                          np.diff(self.outflow_ptr))
        return reaches, flat, group


def compress_ex(rows, cols, gauges, ddr_match: bool = False, parent_offset=None) -> Compressed:
    """Unioned CONUS COO + gauge rows -> a compressed batch with outflow_idx and
    the compressed parent map (data/collate.rs:197-349). parent_offset: the
    CONUS store's reach-subdivision map (ConusAdjacencyStore.parent_offset) or
    None."""
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    gauges = np.ascontiguousarray(gauges, np.int32)
    G = gauges.shape[0]
    cap = 2 * rows.shape[0] + G + 1
    active = np.zeros(cap, np.int32)
    orow = np.zeros(max(rows.shape[0], 1), np.int32)
    ocol = np.zeros(max(rows.shape[0], 1), np.int32)
    og = np.zeros(max(G, 1), np.int32)
    po = None if parent_offset is None else np.ascontiguousarray(parent_offset, np.int32)
    npar_in = 0 if po is None else po.shape[0] - 1
    opo = np.zeros(cap + 1, np.int32)
    optr = np.zeros(G + 1, np.int32)
    ocap = rows.shape[0] + G + 1
    oflow = np.zeros(ocap, np.int32)
    na, ne, npar, nflow = (ctypes.c_int32() for _ in range(4))
    N.check(N.lib.ddrs_compress_ex(rows.shape[0], _ptr(rows), _ptr(cols), G, _ptr(gauges),
                                   int(bool(ddr_match)), npar_in, _ptr(po), _ptr(active), cap,
                                   ctypes.byref(na), _ptr(orow), _ptr(ocol), ctypes.byref(ne),
                                   _ptr(og), _ptr(opo), ctypes.byref(npar), _ptr(optr),
                                   _ptr(oflow), ocap, ctypes.byref(nflow)), "ddrs_compress_ex")
    return Compressed(active[:na.value].copy(), orow[:ne.value].copy(), ocol[:ne.value].copy(),
                      og[:G].copy(), optr, oflow[:nflow.value].copy(),
                      None if po is None else opo[:npar.value + 1].copy())

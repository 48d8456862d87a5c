"""Network stores on disk: the reference's binsparse-COO zarr v3 stores
(src/data/store/zarr.rs:47-300), read through the library's C reader
(ddrs_amd/csrc/ddrs_zarr.cpp, zstd by the system libzstd).

ConusAdjacencyStore and GagesAdjacencyStore mirror the reference's types: the
same fields, the identity parent map for stores written before reach
subdivision, and the same consistency checks (which raise StoreError).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N

_DTYPES = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.uint16, 5: np.int32, 6: np.uint32,
           7: np.int64, 8: np.uint64, 9: np.float32, 10: np.float64}
_CODES = {np.dtype(v): k for k, v in _DTYPES.items()}


class StoreError(ValueError):
    """A store that is missing arrays or whose arrays disagree."""


def read_array(array_dir: str, dtype=None) -> np.ndarray:
    """A whole zarr v3 array as numpy (C order). With `dtype`, the stored type
    must be that type, as the reference's typed reads require."""
    path = os.fsencode(array_dir)
    code, ndim, count = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    shape = (ctypes.c_int64 * 32)()
    N.check(N.lib.ddrs_zarr_array_info(path, ctypes.byref(code), ctypes.byref(ndim), shape, 32,
                                       ctypes.byref(count)), "ddrs_zarr_array_info")
    stored = np.dtype(_DTYPES[code.value])
    if dtype is not None and np.dtype(dtype) != stored:
        raise StoreError(f"{array_dir} holds {stored}, not {np.dtype(dtype)}")
    out = np.empty(tuple(shape[: ndim.value]), stored)
    N.check(N.lib.ddrs_zarr_read_array(path, _CODES[stored], out.ctypes.data if out.size else None,
                                       int(count.value)), "ddrs_zarr_read_array")
    return out


def group_attr(group_dir: str, key: str):
    """A group attribute: a float for numbers, else its text."""
    num = ctypes.c_double()
    text = ctypes.create_string_buffer(512)
    N.check(N.lib.ddrs_zarr_group_attr(os.fsencode(group_dir), key.encode(), ctypes.byref(num),
                                       text, 512), "ddrs_zarr_group_attr")
    return num.value if not np.isnan(num.value) else text.value.decode()


def _try_array(path: str, dtype) -> Optional[np.ndarray]:
    if not os.path.exists(os.path.join(path, "zarr.json")):
        return None
    try:
        return read_array(path, dtype)
    except (N.DdrsError, StoreError):
        return None


@dataclass
class ConusAdjacencyStore:
    """The CONUS-wide network (src/data/store/zarr.rs:47-176): COMIDs in
    topological order, COO rows (downstream) / columns (upstream) in that
    position space, per-reach length_m and slope, and the parent map of a
    subdivided store (identity when the store has none)."""

    path: str
    order: np.ndarray
    parent_order: np.ndarray
    parent_offset: np.ndarray
    length_m: np.ndarray
    slope: np.ndarray
    indices_0: np.ndarray
    indices_1: np.ndarray

    @property
    def n(self) -> int:
        return int(self.order.size)

    @property
    def nnz(self) -> int:
        return int(self.indices_0.size)

    def n_parent(self) -> int:
        return int(self.parent_order.size)

    def outlet_row(self, parent: int) -> int:
        """Last (most downstream) sub-reach row of `parent`."""
        return int(self.parent_offset[parent + 1]) - 1

    @classmethod
    def open(cls, path: str) -> "ConusAdjacencyStore":
        j = lambda a: os.path.join(path, a)
        if not os.path.exists(j("zarr.json")):
            raise StoreError(f"{path}: not a zarr store")
        order = read_array(j("order"), np.int32)
        n = order.size
        parent_order = _try_array(j("parent_order"), np.int32)
        parent_order = order.copy() if parent_order is None else parent_order
        parent_offset = _try_array(j("parent_offset"), np.int32)
        parent_offset = (np.arange(n + 1, dtype=np.int32) if parent_offset is None
                         else parent_offset)
        if parent_offset.size != parent_order.size + 1:
            raise StoreError(f"{path}: parent_offset must have parent_order.len() + 1 entries: "
                             f"{parent_offset.size} vs {parent_order.size + 1}")
        if parent_offset.size == 0 or parent_offset[0] != 0 or parent_offset[-1] != n:
            raise StoreError(f"{path}: parent_offset must run 0..{n}")
        length_m = read_array(j("length_m"), np.float32)
        slope = read_array(j("slope"), np.float32)
        if length_m.size != n or slope.size != n:
            raise StoreError(f"{path}: order/length_m/slope lengths disagree: {n} / "
                             f"{length_m.size} / {slope.size}")
        i0 = read_array(j("indices_0"), np.int32)
        i1 = read_array(j("indices_1"), np.int32)
        if i0.size != i1.size:
            raise StoreError(f"{path}: indices_0 / indices_1 length mismatch: {i0.size} vs "
                             f"{i1.size}")
        return cls(path, order, parent_order, parent_offset, length_m, slope, i0, i1)


@dataclass
class GageSubgraph:
    """One gauge's upstream subgraph in CONUS position space
    (src/data/store/zarr.rs:178-226)."""

    staid: str
    gage_idx: int
    gage_catchment: str
    indices_0: np.ndarray
    indices_1: np.ndarray

    def is_headwater(self) -> bool:
        return self.indices_0.size == 0

    def upstream_comids(self, conus: ConusAdjacencyStore) -> np.ndarray:
        pos = np.unique(np.concatenate([self.indices_0, self.indices_1]))
        return conus.order[pos]


class GagesAdjacencyStore:
    """Per-STAID subgraphs (src/data/store/zarr.rs:228-300): only the requested
    gauges are loaded; a STAID without a group is skipped."""

    def __init__(self, path: str, subgraphs: Dict[str, GageSubgraph]):
        self.path = path
        self.subgraphs = subgraphs

    @classmethod
    def open(cls, path: str, staids: Sequence[str]) -> "GagesAdjacencyStore":
        if not os.path.exists(os.path.join(path, "zarr.json")):
            raise StoreError(f"{path}: not a zarr store")
        subs = {}
        for staid in staids:
            g = os.path.join(path, staid)
            if not os.path.exists(os.path.join(g, "zarr.json")):
                continue
            i0 = read_array(os.path.join(g, "indices_0"), np.int32)
            i1 = read_array(os.path.join(g, "indices_1"), np.int32)
            try:
                gi = group_attr(g, "gage_idx")
            except N.DdrsError as e:
                raise StoreError(f"missing or non-integer 'gage_idx' on /{staid}") from e
            if not isinstance(gi, float) or gi != int(gi):
                raise StoreError(f"missing or non-integer 'gage_idx' on /{staid}")
            try:
                gc = group_attr(g, "gage_catchment")
                gc = str(int(gc)) if isinstance(gc, float) and gc == int(gc) else str(gc)
            except N.DdrsError:
                gc = ""
            subs[staid] = GageSubgraph(staid, int(gi), gc, i0, i1)
        return cls(path, subs)

    def get(self, staid: str) -> Optional[GageSubgraph]:
        return self.subgraphs.get(staid)

    def __len__(self) -> int:
        return len(self.subgraphs)


@dataclass
class UnionedCoo:
    """The per-batch union of gauge subgraphs (data/collate.rs:16-25): distinct
    edges in CONUS positions sorted by (row, col), and one (staid, gage_idx,
    gage_catchment) per gauge present in the store, in request order."""

    rows: np.ndarray
    cols: np.ndarray
    gauges: list

    def gauge_rows(self) -> np.ndarray:
        return np.asarray([g[1] for g in self.gauges], np.int32)


def union_subgraphs(staids: Sequence[str], gages: GagesAdjacencyStore) -> UnionedCoo:
    """union_subgraphs (data/collate.rs:31-48): gauges missing from the store
    are skipped; the edges of the present ones are unioned, deduplicated and
    sorted by (row, col), the order of the reference's BTreeSet."""
    gauges, r, c = [], [], []
    for s in staids:
        g = gages.get(s)
        if g is None:
            continue
        gauges.append((s, int(g.gage_idx), g.gage_catchment))
        r.append(np.asarray(g.indices_0, np.int64))
        c.append(np.asarray(g.indices_1, np.int64))
    if r:
        key = np.unique(np.concatenate(r) * (1 << 32) + np.concatenate(c))
        rows, cols = (key >> 32).astype(np.int32), (key & 0xFFFFFFFF).astype(np.int32)
    else:
        rows = cols = np.zeros(0, np.int32)
    return UnionedCoo(rows, cols, gauges)

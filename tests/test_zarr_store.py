"""The zarr v3 network-store reader (ddrs_zarr_* through ddrs_amd.store) on
the CPU: the reference's Juniata stores (copied to tests/golden/juniata by
tests/golden/make_juniata_fixture.py) against an independent decode (json +
pyarrow's zstd + numpy), synthetic stores for every supported layout, and the
reference's consistency errors (src/data/store/zarr.rs:80-160)."""
import json
import os

import numpy as np
import pyarrow as pa
import pytest

from ddrs_amd import _native as N
from ddrs_amd import store as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "juniata")
CONUS = os.path.join(GOLD, "juniata_conus_adjacency.zarr")
GAGES = os.path.join(GOLD, "juniata_gages_adjacency.zarr")


def independent_read(array_dir):
    """Single-chunk bytes + zstd arrays, decoded without the library."""
    m = json.load(open(os.path.join(array_dir, "zarr.json")))
    dt = np.dtype(m["data_type"])
    n = int(np.prod(m["shape"]))
    raw = open(os.path.join(array_dir, "c", "0"), "rb").read()
    return np.frombuffer(pa.decompress(raw, decompressed_size=n * dt.itemsize, codec="zstd"), dt)


def test_juniata_conus_store_matches_independent_decode():
    st = S.ConusAdjacencyStore.open(CONUS)
    for name in ("order", "indices_0", "indices_1", "length_m", "slope"):
        np.testing.assert_array_equal(getattr(st, name), independent_read(os.path.join(CONUS, name)))
    assert st.n == 213 and st.nnz == 212
    # no parent map in this store: the identity
    np.testing.assert_array_equal(st.parent_order, st.order)
    np.testing.assert_array_equal(st.parent_offset, np.arange(214))
    assert st.outlet_row(5) == 5
    # a lower-triangular tree (downstream row after upstream column), every
    # reach but the outlet drains once
    assert (st.indices_0 > st.indices_1).all()
    assert np.bincount(st.indices_1, minlength=213).max() == 1
    assert (st.length_m > 0).all() and (st.slope >= 0).all()
    # uint8 values are not read as int32
    with pytest.raises(S.StoreError):
        S.read_array(os.path.join(CONUS, "values"), np.int32)
    assert (S.read_array(os.path.join(CONUS, "values")) == 1).all()


def test_juniata_gage_store():
    conus = S.ConusAdjacencyStore.open(CONUS)
    gs = S.GagesAdjacencyStore.open(GAGES, ["01567000", "99999999"])
    assert len(gs) == 1 and gs.get("99999999") is None
    g = gs.get("01567000")
    assert g.gage_idx == 212 and g.gage_catchment == "73005278"
    assert not g.is_headwater()
    np.testing.assert_array_equal(g.indices_0, independent_read(os.path.join(GAGES, "01567000", "indices_0")))
    # the gauge sits at the network's outlet and its subgraph is the whole network
    assert conus.order[g.gage_idx] == 73005278
    assert set(g.upstream_comids(conus)) == set(conus.order)


def write_array(path, data, chunks, codecs=("bytes", "zstd"), endian="little",
                key="default", sep="/", fill=0, skip=()):
    """A zarr v3 array written the way the stores are (test-side writer)."""
    os.makedirs(path, exist_ok=True)
    data = np.asarray(data)
    cl = []
    for c in codecs:
        if c == "bytes":
            cl.append({"name": "bytes", "configuration": {"endian": endian}})
        else:
            cl.append({"name": c, "configuration": {"level": 0, "checksum": False}})
    meta = {"zarr_format": 3, "node_type": "array", "shape": list(data.shape),
            "data_type": str(data.dtype), "chunk_grid": {"name": "regular", "configuration":
                                                         {"chunk_shape": list(chunks)}},
            "chunk_key_encoding": {"name": key, "configuration": {"separator": sep}},
            "fill_value": fill, "codecs": cl, "attributes": {}}
    json.dump(meta, open(os.path.join(path, "zarr.json"), "w"))
    grid = [range(0, s, c) for s, c in zip(data.shape, chunks)]
    for idx in np.ndindex(*[len(g) for g in grid]):
        if idx in skip:
            continue
        block = np.full(chunks, fill, data.dtype)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, data.shape))
        block[tuple(slice(0, x.stop - x.start) for x in sl)] = data[sl]
        raw = block.astype(block.dtype.newbyteorder(">" if endian == "big" else "<")).tobytes()
        if "zstd" in codecs:
            raw = pa.compress(raw, codec="zstd").to_pybytes()
        parts = [str(i) for i in idx]
        rel = ("c" + sep + sep.join(parts)) if key == "default" else ".".join(parts)
        f = os.path.join(path, *rel.split("/"))
        os.makedirs(os.path.dirname(f), exist_ok=True)
        open(f, "wb").write(raw)


def test_layouts(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.integers(-1000, 1000, 1000).astype(np.int32)
    write_array(tmp_path / "multi", a, [300])
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "multi"), np.int32), a)
    # a missing chunk reads as the fill value
    write_array(tmp_path / "missing", a, [300], fill=-7, skip={(1,)})
    want = a.copy()
    want[300:600] = -7
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "missing")), want)
    b = rng.normal(size=(37, 21)).astype(np.float64)
    write_array(tmp_path / "twod", b, [10, 8])
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "twod")), b)
    c = rng.normal(size=77).astype(np.float32)
    write_array(tmp_path / "big", c, [20], endian="big")
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "big")), c)
    write_array(tmp_path / "raw", c, [50], codecs=("bytes",))
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "raw")), c)
    write_array(tmp_path / "v2", b, [16, 16], key="v2", sep=".")
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "v2")), b)
    d = np.arange(10, dtype=np.uint64) * 2**40
    write_array(tmp_path / "u64", d, [4])
    np.testing.assert_array_equal(S.read_array(str(tmp_path / "u64")), d)


def test_errors(tmp_path):
    c = np.arange(30, dtype=np.float32)
    write_array(tmp_path / "gz", c, [10], codecs=("bytes", "gzip"))
    with pytest.raises(N.DdrsError, match="codec"):
        S.read_array(str(tmp_path / "gz"))
    write_array(tmp_path / "f", c, [10])
    with pytest.raises(S.StoreError):
        S.read_array(str(tmp_path / "f"), np.int32)
    # a corrupt chunk
    open(tmp_path / "f" / "c" / "1", "wb").write(b"not zstd")
    with pytest.raises(N.DdrsError, match="zstd"):
        S.read_array(str(tmp_path / "f"))
    with pytest.raises(N.DdrsError):
        S.read_array(str(tmp_path / "nowhere"))
    # store-level checks
    root = tmp_path / "store"
    os.makedirs(root)
    json.dump({"zarr_format": 3, "node_type": "group", "attributes": {}},
              open(root / "zarr.json", "w"))
    write_array(root / "order", np.arange(5, dtype=np.int32), [5])
    write_array(root / "length_m", np.ones(5, np.float32), [5])
    write_array(root / "slope", np.ones(4, np.float32), [4])
    write_array(root / "indices_0", np.array([1, 2], np.int32), [2])
    write_array(root / "indices_1", np.array([0, 1], np.int32), [2])
    with pytest.raises(S.StoreError, match="lengths disagree"):
        S.ConusAdjacencyStore.open(str(root))
    write_array(root / "slope", np.ones(5, np.float32), [5])
    st = S.ConusAdjacencyStore.open(str(root))
    assert st.n == 5 and st.nnz == 2
    write_array(root / "parent_order", np.arange(3, dtype=np.int32), [3])
    write_array(root / "parent_offset", np.array([0, 2, 4, 4], np.int32), [4])
    with pytest.raises(S.StoreError, match="0..5"):
        S.ConusAdjacencyStore.open(str(root))
    write_array(root / "parent_offset", np.array([0, 2, 4, 5], np.int32), [4])
    st = S.ConusAdjacencyStore.open(str(root))
    assert st.n_parent() == 3 and st.outlet_row(1) == 3

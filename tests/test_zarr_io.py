"""Zarr v2 chunk I/O: stores written here read back exactly (raw and zlib
chunks, F and C chunk order, '.' and '/' separators, partial edge chunks),
regions across chunk boundaries equal numpy slices of the array, absent chunks
read as the fill value, gzip chunks decode, and blosc is refused clearly."""
import gzip
import json
import os

import numpy as np
import pytest

from petakit5d_amd.zarr_io import read_zarr, write_zarr


def _arr(shape, dtype=np.uint16, seed=0):
    g = np.random.default_rng(seed)
    return np.asfortranarray(g.integers(0, 60000, size=shape).astype(dtype))


@pytest.mark.parametrize("comp", [None, "zlib"])
@pytest.mark.parametrize("order", ["F", "C"])
@pytest.mark.parametrize("sep", [".", "/"])
def test_round_trip(tmp_path, comp, order, sep):
    a = _arr((37, 29, 11))
    p = str(tmp_path / "s.zarr")
    write_zarr(p, a, chunks=(16, 8, 4), compressor=comp, order=order, dimension_separator=sep)
    b = read_zarr(p)
    assert b.flags["F_CONTIGUOUS"] and b.dtype == a.dtype
    np.testing.assert_array_equal(b, a)


def test_regions_across_chunk_boundaries(tmp_path):
    a = _arr((50, 44, 30), np.float32, seed=2)
    p = str(tmp_path / "f.zarr")
    write_zarr(p, a, chunks=(16, 12, 7))
    for bbox in ([1, 1, 1, 50, 44, 30], [3, 5, 2, 47, 40, 29], [16, 12, 7, 17, 13, 8], [50, 44, 30, 50, 44, 30]):
        got = read_zarr(p, bbox)
        np.testing.assert_array_equal(got, a[bbox[0] - 1:bbox[3], bbox[1] - 1:bbox[4], bbox[2] - 1:bbox[5]])
    with pytest.raises(ValueError):
        read_zarr(p, [0, 1, 1, 5, 5, 5])


def test_absent_chunks_read_fill_value_and_gzip(tmp_path):
    a = _arr((20, 20, 8))
    p = str(tmp_path / "g.zarr")
    write_zarr(p, a, chunks=(10, 10, 8), compressor=None, fill_value=7)
    os.remove(os.path.join(p, "1.1.0"))
    b = read_zarr(p)
    np.testing.assert_array_equal(b[:10], a[:10])
    assert np.all(b[10:, 10:] == 7)
    # the same store with gzip chunks (the reference reader's other codec)
    m = json.load(open(os.path.join(p, ".zarray")))
    m["compressor"] = {"id": "gzip", "level": 1}
    json.dump(m, open(os.path.join(p, ".zarray"), "w"))
    for fn in ("0.0.0", "0.1.0", "1.0.0"):
        raw = open(os.path.join(p, fn), "rb").read()
        open(os.path.join(p, fn), "wb").write(gzip.compress(raw))
    np.testing.assert_array_equal(read_zarr(p), b)


def test_blosc_refused(tmp_path):
    a = _arr((8, 8, 8))
    p = str(tmp_path / "b.zarr")
    write_zarr(p, a, chunks=(8, 8, 8))
    m = json.load(open(os.path.join(p, ".zarray")))
    m["compressor"] = {"id": "blosc", "cname": "lz4", "clevel": 5, "shuffle": 1, "blocksize": 0}
    json.dump(m, open(os.path.join(p, ".zarray"), "w"))
    with pytest.raises(NotImplementedError):
        read_zarr(p)

"""Zarr v2 chunk I/O: stores written here read back exactly (raw and zlib
chunks, F and C chunk order, '.' and '/' separators, partial edge chunks),
regions across chunk boundaries equal numpy slices of the array, absent chunks
read as the fill value, gzip chunks decode; blosc chunks (lz4 streams, byte
shuffle) round-trip, decode when assembled by hand from the published format,
and the LZ4 stream codec agrees with the system's liblz4 both ways."""
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


def test_blosc_round_trip(tmp_path):
    """Stores written with blosc lz4 + byte shuffle (the reference writer's
    default) read back exactly: several dtypes, chunks whose size leaves a
    short last block, and incompressible data (stored as a plain copy)."""
    g = np.random.default_rng(3)
    for dt in (np.uint16, np.float32, np.uint8):
        a = (g.random((70, 66, 40)) * 200).astype(dt)
        p = str(tmp_path / f"b_{np.dtype(dt).name}.zarr")
        write_zarr(p, a, chunks=(64, 64, 32), compressor="blosc")
        m = json.load(open(os.path.join(p, ".zarray")))
        assert m["compressor"] == {"id": "blosc", "cname": "lz4", "clevel": 5, "shuffle": 1, "blocksize": 0}
        np.testing.assert_array_equal(read_zarr(p), a)
        np.testing.assert_array_equal(read_zarr(p, [5, 60, 3, 70, 66, 40]), a[4:70, 59:66, 2:40])
    noise = g.integers(0, 256, size=(40, 30, 20), dtype=np.uint8)
    p = str(tmp_path / "noise.zarr")
    write_zarr(p, noise, chunks=(40, 30, 20), compressor="blosc")
    assert open(os.path.join(p, "0.0.0"), "rb").read()[2] & 0x2  # plain copy
    np.testing.assert_array_equal(read_zarr(p), noise)


def _liblz4():
    import ctypes
    try:
        L = ctypes.CDLL("liblz4.so.1")
    except OSError:
        pytest.skip("liblz4 is not installed")
    L.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    return L


def _lz4_ref_compress(L, b):
    import ctypes
    out = ctypes.create_string_buffer(len(b) + len(b) // 255 + 64)
    n = L.LZ4_compress_default(b, out, len(b), len(out))
    assert n > 0 or len(b) == 0
    return out.raw[:n]


@pytest.mark.parametrize("kind", ["random", "runs", "image", "tiny"])
def test_lz4_against_liblz4(kind):
    """The LZ4 block codec against the system's liblz4: blocks compressed by
    liblz4 decode here, blocks compressed here decode with liblz4."""
    import ctypes
    from petakit5d_amd import zarr_io as Z
    L = _liblz4()
    g = np.random.default_rng(8)
    cases = {
        "random": [g.integers(0, 256, size=200_000, dtype=np.uint8).tobytes()],
        "runs": [bytes(np.repeat(g.integers(0, 3, size=3000, dtype=np.uint8), g.integers(1, 400, size=3000)))],
        "image": [(g.integers(0, 3000, size=150_000) // 9 * 9).astype(np.uint16).tobytes()],
        "tiny": [bytes(range(n)) for n in range(0, 20)] + [b"a" * n for n in range(1, 40)],
    }[kind]
    P = Z._lz4()
    for b in cases:
        ref = _lz4_ref_compress(L, b)
        if b:
            assert Z._lz4_decode(P, ref, len(b)) == b
        mine = Z._lz4_encode(P, b)
        out = ctypes.create_string_buffer(max(len(b), 1))
        n = L.LZ4_decompress_safe(mine, out, len(mine), len(b))
        assert n == len(b) and out.raw[:n] == b


def test_blosc_chunk_assembled_from_the_format(tmp_path):
    """A blosc chunk assembled by hand from the published format with
    liblz4 streams (byte shuffle, split streams, a raw stream, a short last
    block that is not split) decodes to the original bytes."""
    import struct
    from petakit5d_amd.zarr_io import blosc_decode
    L = _liblz4()
    g = np.random.default_rng(9)
    a = (g.integers(0, 500, size=3000) // 7 * 7).astype(np.uint16)
    data = a.tobytes()  # 6000 bytes
    bs = 2048  # blocks of 2048, 2048 and 1904 bytes
    body, starts, off = [], [], 16 + 4 * 3
    for j in range(3):
        blk = np.frombuffer(data[j * bs:(j + 1) * bs], dtype=np.uint8)
        sh = blk.reshape(-1, 2).T.reshape(-1).tobytes()  # byte shuffle
        streams = [sh[:len(sh) // 2], sh[len(sh) // 2:]] if len(blk) == bs else [sh]
        starts.append(off)
        for k, st in enumerate(streams):
            c = st if (j == 1 and k == 1) else _lz4_ref_compress(L, st)  # one stream stored raw
            body.append(struct.pack("<i", len(c)) + c)
            off += 4 + len(c)
    chunk = struct.pack("<BBBBIII", 2, 1, 0x1 | (1 << 5), 2, len(data), bs, off) + \
        struct.pack("<3i", *starts) + b"".join(body)
    assert blosc_decode(chunk) == data


def test_blosc_unsupported_codecs_refused(tmp_path):
    import struct
    from petakit5d_amd.zarr_io import blosc_decode
    for flags in ((4 << 5), (0 << 5), (1 << 5) | 0x4):  # zstd, blosclz, bit shuffle
        chunk = struct.pack("<BBBBIII", 2, 1, flags, 2, 64, 64, 100) + b"\0" * 84
        with pytest.raises(NotImplementedError):
            blosc_decode(chunk)

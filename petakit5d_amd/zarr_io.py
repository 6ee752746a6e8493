"""Zarr v2 chunk I/O for the large-volume path: read a region of a store into
a column-major array, write an array as a store, and deconvolve a store
through the chunked driver.

Metadata follow the reference's reader and writer (io/cpp-zarr/src/zarr.cpp:
16-18 defaults: dtype '<u2', order 'F', dimension_separator '.', fill_value 0;
:51-96 parsing: chunks, compressor, dimension_separator '.' or '/', fill_value
null / NaN -> 0).  Arrays use MATLAB's size order (shape [y x z], dim 1
contiguous).  Compressors: none, zlib, gzip and blosc (the reference's default:
lz4, byte shuffle, clevel 5, zarr.cpp:15-18) with its lz4 / lz4hc or zlib
streams; blosc's blosclz, snappy and zstd streams and bit shuffle raise
NotImplementedError.  Edge chunks are stored full size, as the Zarr v2 format
requires.

Blosc chunks follow the published blosc 1 format: a 16-byte header (format
versions, flags, type size, uncompressed bytes, block size, compressed bytes),
then, unless the flags say the chunk is a plain copy, one 32-bit start offset
per block; a block holds type-size streams (one stream when the flags forbid
splitting, when the type size exceeds 16, when a stream would be shorter than
128 bytes, and for the last, short block), each a 32-bit length and that many
bytes, stored raw when the length equals the stream's size.  Byte shuffle
groups byte k of every element of a block together.  The LZ4 streams go
through libpkdecon.so's host codec (pk_lz4_decode / pk_lz4_encode)."""
from __future__ import annotations

import ctypes
import gzip
import json
import os
import struct
import zlib

import numpy as np

__all__ = ["read_zarr", "write_zarr", "decon_zarr", "blosc_decode", "blosc_encode"]

_BLOSC_MAX_SPLITS, _BLOSC_MIN_STREAM = 16, 128
_BLOSC_CODES = {"blosclz": 0, "lz4": 1, "lz4hc": 1, "snappy": 2, "zlib": 3, "zstd": 4}


def _lz4():
    from ._lib import load
    L = load()
    for name in ("pk_lz4_decode", "pk_lz4_encode"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                      ctypes.POINTER(ctypes.c_size_t)]
    L.pk_lz4_encode_bound.restype = ctypes.c_size_t
    L.pk_lz4_encode_bound.argtypes = [ctypes.c_size_t]
    return L


def _lz4_decode(L, src: bytes, size: int) -> bytes:
    a = np.frombuffer(src, dtype=np.uint8)
    out = np.empty(max(size, 1), dtype=np.uint8)
    n = ctypes.c_size_t(0)
    rc = L.pk_lz4_decode(a.ctypes.data if a.size else None, a.size, out.ctypes.data, size, ctypes.byref(n))
    if rc != 0 or n.value != size:
        raise ValueError("corrupt lz4 stream in a blosc chunk")
    return out[:size].tobytes()


def _lz4_encode(L, src: bytes) -> bytes:
    a = np.frombuffer(src, dtype=np.uint8)
    cap = int(L.pk_lz4_encode_bound(a.size))
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    if L.pk_lz4_encode(a.ctypes.data if a.size else None, a.size, out.ctypes.data, cap, ctypes.byref(n)) != 0:
        raise RuntimeError("pk_lz4_encode failed")
    return out[:n.value].tobytes()


def _splits(flags, typesize, bsize, last_short):
    if (flags & 0x10) or last_short or typesize > _BLOSC_MAX_SPLITS or bsize // typesize < _BLOSC_MIN_STREAM:
        return 1
    return typesize


def _unshuffle(b: bytes, typesize: int) -> bytes:
    q = len(b) // typesize
    if typesize <= 1 or q == 0:
        return b
    a = np.frombuffer(b, dtype=np.uint8)
    body = a[:q * typesize].reshape(typesize, q).T.reshape(-1)
    return body.tobytes() + a[q * typesize:].tobytes()


def _shuffle(b: bytes, typesize: int) -> bytes:
    q = len(b) // typesize
    if typesize <= 1 or q == 0:
        return b
    a = np.frombuffer(b, dtype=np.uint8)
    body = a[:q * typesize].reshape(q, typesize).T.reshape(-1)
    return body.tobytes() + a[q * typesize:].tobytes()


def blosc_decode(raw: bytes) -> bytes:
    """One blosc 1 chunk -> its bytes."""
    if len(raw) < 16:
        raise ValueError("blosc chunk shorter than its header")
    flags, typesize = raw[2], max(raw[3], 1)
    nbytes, blocksize, cbytes = struct.unpack_from("<III", raw, 4)
    if flags & 0x2:  # plain copy
        return bytes(raw[16:16 + nbytes])
    if flags & 0x4:
        raise NotImplementedError("blosc bit shuffle is not supported by petakit5d_amd")
    code = flags >> 5
    if code not in (1, 3):
        name = {0: "blosclz", 2: "snappy", 4: "zstd"}.get(code, str(code))
        raise NotImplementedError(f"blosc '{name}' streams are not supported by petakit5d_amd (lz4, lz4hc, zlib are)")
    if nbytes == 0:
        return b""
    if blocksize == 0:
        raise ValueError("blosc chunk with a zero block size")
    L = _lz4() if code == 1 else None
    nblocks = -(-nbytes // blocksize)
    starts = struct.unpack_from(f"<{nblocks}i", raw, 16)
    out = []
    for j, pos in enumerate(starts):
        bsize = min(blocksize, nbytes - j * blocksize)
        ns = _splits(flags, typesize, bsize, bsize < blocksize)
        neb = bsize // ns
        parts = []
        for _ in range(ns):
            (cs,) = struct.unpack_from("<i", raw, pos)
            pos += 4
            seg = raw[pos:pos + cs]
            pos += cs
            if cs == neb:
                parts.append(bytes(seg))
            elif code == 1:
                parts.append(_lz4_decode(L, seg, neb))
            else:
                parts.append(zlib.decompress(seg))
        blk = b"".join(parts)
        if len(blk) != bsize:
            raise ValueError("corrupt blosc block")
        out.append(_unshuffle(blk, typesize) if flags & 0x1 else blk)
    return b"".join(out)


def blosc_encode(data: bytes, typesize: int, shuffle: int = 1, blocksize: int = 1 << 18) -> bytes:
    """Bytes -> one blosc 1 chunk with lz4 streams (byte shuffle when shuffle=1)."""
    if shuffle not in (0, 1):
        raise NotImplementedError("blosc bit shuffle is not supported by petakit5d_amd")
    L = _lz4()
    nbytes = len(data)
    typesize = max(1, int(typesize))
    bsz = max(typesize, min(blocksize, nbytes) // typesize * typesize) if nbytes else 0
    flags = (1 << 5) | (0x1 if shuffle and typesize > 1 else 0)
    nblocks = -(-nbytes // bsz) if nbytes else 0
    body, starts = [], []
    off = 16 + 4 * nblocks
    for j in range(nblocks):
        blk = data[j * bsz:(j + 1) * bsz]
        if flags & 0x1:
            blk = _shuffle(blk, typesize)
        ns = _splits(flags, typesize, len(blk), len(blk) < bsz)
        neb = len(blk) // ns
        starts.append(off)
        for k in range(ns):
            seg = blk[k * neb:(k + 1) * neb]
            c = _lz4_encode(L, seg)
            if len(c) >= neb:
                c = seg
            body.append(struct.pack("<i", len(c)) + c)
            off += 4 + len(c)
    if off > 16 + nbytes:  # incompressible: a plain copy
        return struct.pack("<BBBBIII", 2, 1, 0x2 | (flags & 0xE0), min(typesize, 255), nbytes, bsz, 16 + nbytes) + data
    head = struct.pack("<BBBBIII", 2, 1, flags, min(typesize, 255), nbytes, bsz, off)
    return head + struct.pack(f"<{nblocks}i", *starts) + b"".join(body)


def _meta(path):
    with open(os.path.join(path, ".zarray")) as f:
        m = json.load(f)
    if m.get("zarr_format", 2) != 2:
        raise ValueError("only Zarr v2 stores are supported")
    sep = m.get("dimension_separator") or "."
    if sep not in (".", "/"):
        raise ValueError("metadataIncomplete:dimension_separatorIncorrect")
    fv = m.get("fill_value", 0)
    if fv is None or fv in ("NaN", "null"):
        fv = 0
    return m, sep, fv


def _decode(raw, comp):
    if comp is None:
        return raw
    cid = comp.get("id")
    if cid == "zlib":
        return zlib.decompress(raw)
    if cid == "gzip":
        return gzip.decompress(raw)
    if cid == "blosc":
        return blosc_decode(raw)
    raise NotImplementedError(f"chunk compressor '{cid}' is not supported by petakit5d_amd (none, zlib, gzip, blosc are)")


def _chunk_path(path, idx, sep):
    return os.path.join(path, sep.join(str(i) for i in idx)) if sep == "." else os.path.join(path, *map(str, idx))


def read_zarr(path, bbox=None):
    """Region bbox (1-based inclusive [y0 x0 z0 y1 x1 z1]; the whole array
    when None) of the store at path, as a column-major numpy array.  Chunks
    that are absent read as the fill value."""
    m, sep, fv = _meta(path)
    shape = [int(n) for n in m["shape"]]
    chunks = [int(n) for n in m["chunks"]]
    dt = np.dtype(m["dtype"])
    order = m.get("order", "C")
    if bbox is None:
        lo, hi = [0] * len(shape), shape
    else:
        b = [int(v) for v in bbox]
        nd = len(shape)
        lo, hi = [b[i] - 1 for i in range(nd)], [b[nd + i] for i in range(nd)]
        if any(lo[i] < 0 or hi[i] > shape[i] or lo[i] >= hi[i] for i in range(nd)):
            raise ValueError("bbox outside the array")
    out = np.full([h - l for l, h in zip(lo, hi)], fv, dtype=dt.newbyteorder("="), order="F")
    first = [l // c for l, c in zip(lo, chunks)]
    last = [(h - 1) // c for h, c in zip(hi, chunks)]
    for idx in np.ndindex(*[b - a + 1 for a, b in zip(first, last)]):
        ci = [a + i for a, i in zip(first, idx)]
        fn = _chunk_path(path, ci, sep)
        if not os.path.exists(fn):
            continue
        with open(fn, "rb") as f:
            raw = _decode(f.read(), m.get("compressor"))
        blk = np.frombuffer(raw, dtype=dt).reshape(chunks, order=order)
        c0 = [c * k for c, k in zip(ci, chunks)]
        s_lo = [max(l, a) for l, a in zip(lo, c0)]
        s_hi = [min(h, a + k) for h, a, k in zip(hi, c0, chunks)]
        src = tuple(slice(a - c, b - c) for a, b, c in zip(s_lo, s_hi, c0))
        dst = tuple(slice(a - l, b - l) for a, b, l in zip(s_lo, s_hi, lo))
        out[dst] = blk[src]
    return out


def _create_store(path, shape, dtype, chunks, compressor, level, dimension_separator, order, fill_value):
    if dimension_separator not in (".", "/"):
        raise ValueError("dimension_separator must be '.' or '/'")
    if compressor not in (None, "zlib", "blosc"):
        raise NotImplementedError("write_zarr: compressor must be 'zlib', 'blosc' or None")
    chunks = [int(min(c, n)) if n > 0 else int(c) for c, n in zip(chunks, shape)]
    comp = {None: None, "zlib": {"id": "zlib", "level": int(level)},
            "blosc": {"id": "blosc", "cname": "lz4", "clevel": 5, "shuffle": 1, "blocksize": 0}}[compressor]
    os.makedirs(path, exist_ok=True)
    meta = {"zarr_format": 2, "shape": [int(n) for n in shape], "chunks": chunks, "dtype": np.dtype(dtype).str,
            "compressor": comp, "fill_value": fill_value, "order": order, "filters": None,
            "dimension_separator": dimension_separator}
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f)
    return meta


def _write_chunks(path, meta, a, origin=(0, 0, 0)):
    """Write the chunks covered by a, placed at origin (0-based, a multiple
    of the chunk shape); chunks past the array's end are stored full size."""
    chunks, shape = meta["chunks"], meta["shape"]
    comp = meta["compressor"]
    cid = None if comp is None else comp["id"]
    sep, order = meta["dimension_separator"], meta["order"]
    if any(o % c for o, c in zip(origin, chunks)):
        raise ValueError("region origin not on a chunk boundary")
    c_first = [o // c for o, c in zip(origin, chunks)]
    c_count = [-(-n // c) for n, c in zip(a.shape, chunks)]
    for rel in np.ndindex(*c_count):
        ci = [f + r for f, r in zip(c_first, rel)]
        c0 = [r * k for r, k in zip(rel, chunks)]
        src = tuple(slice(s0, min(s0 + k, n)) for s0, k, n in zip(c0, chunks, a.shape))
        g_end = [o + x.stop for o, x in zip(origin, src)]
        if any(e < min((i + 1) * k, n) for e, i, k, n in zip(g_end, ci, chunks, shape)):
            raise ValueError("region ends inside a chunk that is not the array's last")
        blk = np.full(chunks, meta["fill_value"], dtype=a.dtype)
        blk[tuple(slice(0, x.stop - x.start) for x in src)] = a[src]
        raw = blk.tobytes(order=order)
        if cid == "zlib":
            raw = zlib.compress(raw, int(comp["level"]))
        elif cid == "blosc":
            raw = blosc_encode(raw, a.dtype.itemsize)
        fn = _chunk_path(path, ci, sep)
        os.makedirs(os.path.dirname(fn), exist_ok=True)
        with open(fn, "wb") as f:
            f.write(raw)


def write_zarr(path, data, chunks=(256, 256, 256), compressor="zlib", level=1, dimension_separator=".", order="F",
               fill_value=0):
    """Write data (shape = MATLAB size) as a Zarr v2 store: zlib-compressed
    (compressor='zlib'), blosc lz4 with byte shuffle (compressor='blosc', the
    reference writer's default, zarr.cpp:15-18) or raw (None) chunks,
    full-size edge chunks."""
    a = np.asarray(data)
    meta = _create_store(path, a.shape, a.dtype, chunks, compressor, level, dimension_separator, order, fill_value)
    _write_chunks(path, meta, a)


def decon_zarr(in_path, out_path, PSF_f, PSF_b, NUMIT, method="omw", batchSize=(512, 512, 512), chunks=None,
               compressor=None, groupSize=20, flagPath=None, Overwrite=False, **kwargs):
    """Deconvolve the store at in_path into a store at out_path as the Zarr
    large-volume driver does (RLdecon_large_in_place.m:172-266 with
    RLdecon_for_zarr_block.m:70-177), streaming: batches of batchSize
    (rounded up to whole output chunks) with a border of round((size(psf) +
    10) / 2), all of the same size (sameBatchSize = true, clipped at the
    volume's start), sorted by voxel count and then by box (the reference's
    sortrows, :179-182), taken groupSize at a time (taskSize 20); each group
    is deconvolved by ONE batched engine call and each batch's region is
    written straight into the output store.  The next group's batches are
    read from the store while the current group runs on the device, and the
    previous group's regions are written meanwhile.

    flagPath (default out_path + '.zarr_flag', the reference's zarr_flag
    folder): a group that completes writes blocks_<first>_<last>.mat (1-based
    batch numbers in the sorted order, RLdecon_large_in_place.m:220-221,
    RLdecon_for_zarr_block.m:173-177); a later run skips groups whose flag
    file exists unless Overwrite (:79-86), so an interrupted run resumes.
    The output has the input's chunk shape unless chunks is given, and the
    input's compressor (blosc or zlib; zlib for other inputs) unless
    compressor is given.  Returns None."""
    import time
    from concurrent.futures import ThreadPoolExecutor

    from .batch import decon_batch
    from .chunks import XR_zarrChunkCoordinatesExtraction, _vec3, matlab_round
    m, _, _ = _meta(in_path)
    shape = [int(n) for n in m["shape"]]
    if len(shape) != 3:
        raise ValueError("a 3-D store is required")
    if compressor is None:
        cid = (m.get("compressor") or {}).get("id")
        compressor = "blosc" if cid == "blosc" else "zlib"
    ochunks = [int(min(c, n)) for c, n in zip(chunks or m["chunks"], shape)]
    bs = _vec3(batchSize, "batchSize")
    bs = np.array([-(-int(b) // c) * c for b, c in zip(bs, ochunks)])  # whole output chunks per batch
    psf_shape = np.shape(PSF_f)
    border = matlab_round((np.asarray(psf_shape, dtype=np.float64) + 10) / 2)
    batch, region, _ = zarr_batches(shape, bs, border)
    flag_dir = flagPath or (str(out_path).rstrip("/") + ".zarr_flag")
    os.makedirs(flag_dir, exist_ok=True)
    odt = np.uint16 if kwargs.get("save16bit", True) else np.float32
    meta_out = _meta(out_path)[0] if os.path.exists(os.path.join(out_path, ".zarray")) else \
        _create_store(out_path, shape, odt, ochunks, compressor, 1, ".", "F", 0)
    gsz = max(1, int(groupSize))
    groups = []
    for g0 in range(0, len(batch), gsz):
        flag = os.path.join(flag_dir, f"blocks_{g0 + 1}_{min(g0 + gsz, len(batch))}.mat")
        if os.path.exists(flag):
            if not Overwrite:
                continue  # done by an earlier run
            os.remove(flag)
        groups.append((g0, flag))

    def read_group(g0):
        return [read_zarr(in_path, b) for b in batch[g0:g0 + gsz]]

    def write_group(g0, outs, flag, t0):
        for b, r, o in zip(batch[g0:g0 + gsz], region[g0:g0 + gsz], outs):
            lo = r[:3] - b[:3]
            hi = lo + (r[3:] - r[:3]) + 1
            _write_chunks(out_path, meta_out, np.asarray(o)[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]],
                          origin=tuple(int(v) - 1 for v in r[:3]))
        from scipy.io import savemat
        tmp = flag[:-4] + f"_{os.getpid()}.mat"
        savemat(tmp, {"t": time.perf_counter() - t0})
        os.replace(tmp, flag)

    with ThreadPoolExecutor(max_workers=1) as reader, ThreadPoolExecutor(max_workers=1) as writer:
        pending = reader.submit(read_group, groups[0][0]) if groups else None
        writes = []
        for gi, (g0, flag) in enumerate(groups):
            t0 = time.perf_counter()
            ims = pending.result()
            pending = reader.submit(read_group, groups[gi + 1][0]) if gi + 1 < len(groups) else None
            outs = decon_batch(ims, PSF_f, PSF_b, NUMIT, method=method, **kwargs)
            writes.append(writer.submit(write_group, g0, outs, flag, t0))
        for w in writes:
            w.result()


def zarr_batches(shape, batchSize, border):
    """Batch, region and local boxes of the Zarr driver: sameBatchSize = true,
    BlockSize = batchSize, sorted by batch voxel count, then by the box
    itself (sortrows([prod(size), box]), RLdecon_large_in_place.m:174-182)."""
    from .chunks import XR_zarrChunkCoordinatesExtraction
    batch, region, local = XR_zarrChunkCoordinatesExtraction(shape, batchSize=batchSize, BlockSize=batchSize,
                                                             sameBatchSize=True, BorderSize=border)
    vox = np.prod(batch[:, 3:] - batch[:, :3] + 1, axis=1)
    keys = np.concatenate([vox[:, None], batch], axis=1)
    order = np.lexsort(keys.T[::-1])  # sortrows: first column first
    return batch[order], region[order], local[order]

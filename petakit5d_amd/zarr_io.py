"""Zarr v2 chunk I/O for the large-volume path: read a region of a store into
a column-major array, write an array as a store, and deconvolve a store
through the chunked driver.

Metadata follow the reference's reader and writer (io/cpp-zarr/src/zarr.cpp:
16-18 defaults: dtype '<u2', order 'F', dimension_separator '.', fill_value 0;
:51-96 parsing: chunks, compressor, dimension_separator '.' or '/', fill_value
null / NaN -> 0).  Arrays use MATLAB's size order (shape [y x z], dim 1
contiguous).  Compressors: none, zlib and gzip; blosc (the reference's default
codec) is not available in this build and raises NotImplementedError.  Edge
chunks are stored full size, as the Zarr v2 format requires."""
from __future__ import annotations

import gzip
import json
import os
import zlib

import numpy as np

__all__ = ["read_zarr", "write_zarr", "decon_zarr"]


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
    raise NotImplementedError(f"chunk compressor '{cid}' is not available in this build (none, zlib, gzip are)")


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


def write_zarr(path, data, chunks=(256, 256, 256), compressor="zlib", level=1, dimension_separator=".", order="F",
               fill_value=0):
    """Write data (shape = MATLAB size) as a Zarr v2 store: zlib-compressed
    (compressor='zlib') or raw (None) chunks, full-size edge chunks."""
    a = np.asarray(data)
    if dimension_separator not in (".", "/"):
        raise ValueError("dimension_separator must be '.' or '/'")
    chunks = [int(min(c, n)) if n > 0 else int(c) for c, n in zip(chunks, a.shape)]
    comp = None if compressor is None else {"id": "zlib", "level": int(level)}
    if compressor not in (None, "zlib"):
        raise NotImplementedError("write_zarr: compressor must be 'zlib' or None")
    os.makedirs(path, exist_ok=True)
    meta = {"zarr_format": 2, "shape": list(a.shape), "chunks": chunks, "dtype": a.dtype.str,
            "compressor": comp, "fill_value": fill_value, "order": order, "filters": None,
            "dimension_separator": dimension_separator}
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f)
    grid = [-(-n // c) for n, c in zip(a.shape, chunks)]
    for ci in np.ndindex(*grid):
        c0 = [c * k for c, k in zip(ci, chunks)]
        blk = np.full(chunks, fill_value, dtype=a.dtype)
        src = tuple(slice(s, min(s + k, n)) for s, k, n in zip(c0, chunks, a.shape))
        blk[tuple(slice(0, x.stop - x.start) for x in src)] = a[src]
        raw = blk.tobytes(order=order)
        if comp is not None:
            raw = zlib.compress(raw, int(level))
        fn = _chunk_path(path, ci, dimension_separator)
        os.makedirs(os.path.dirname(fn), exist_ok=True)
        with open(fn, "wb") as f:
            f.write(raw)


def decon_zarr(in_path, out_path, PSF_f, PSF_b, NUMIT, method="omw", batchSize=(512, 512, 512), chunks=None,
               **kwargs):
    """Deconvolve the store at in_path into a store at out_path: the volume is
    read, deconvolved in batches of batchSize through one batched engine call
    (decon_large_in_memory), and written with the input's chunk shape unless
    chunks is given."""
    from .chunks import decon_large_in_memory
    m, _, _ = _meta(in_path)
    im = read_zarr(in_path)
    J = decon_large_in_memory(im, PSF_f, PSF_b, NUMIT, method=method, batchSize=batchSize, **kwargs)
    write_zarr(out_path, J, chunks=chunks or m["chunks"])
    return J

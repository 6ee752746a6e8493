"""Batched deconvolution of many chunks of different sizes (one launch per
pass over all chunks), the per-chunk loop of RLdecon_large_in_memory.m:190-260
and RLdecon_for_zarr_block.m:101-171 in one call."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PK_MEM_DEVICE, PK_MEM_HOST, PK_SINGLE, PK_UINT16, PkVol, i64x3
from .decon import _dims3, _is_torch, _opts_struct, _parse, _psf_array


class PkBatchItem(ctypes.Structure):
    _fields_ = [("inp", PkVol), ("out", PkVol)]


def _vol_of(x):
    """(keepalive object, PkVol, kind) for a host numpy array or a column-major CUDA tensor."""
    if _is_torch(x) and x.is_cuda:
        import torch
        d = _dims3(x.shape)
        if x.dtype not in (torch.uint16, torch.float32):
            raise TypeError("chunk class must be uint16 or single")
        if tuple(x.stride()) != (1, d[0], d[0] * d[1]):
            x = x.permute(2, 1, 0).contiguous().permute(2, 1, 0)
        code = PK_UINT16 if x.dtype == torch.uint16 else PK_SINGLE
        return x, PkVol(ctypes.c_void_p(x.data_ptr()), code, PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*d)), "torch"
    a = np.asarray(x)
    if a.dtype not in (np.uint16, np.float32):
        raise TypeError("chunk class must be uint16 or single")
    d = _dims3(a.shape)
    a = np.asfortranarray(a.reshape(d, order="F"))
    code = PK_UINT16 if a.dtype == np.uint16 else PK_SINGLE
    return a, PkVol(ctypes.c_void_p(a.ctypes.data), code, PK_MEM_HOST, (ctypes.c_int64 * 3)(*d)), "numpy"


def decon_batch(chunks, PSF_f, PSF_b, NUMIT, method="omw", **kwargs):
    """Deconvolve every chunk with the same PSF(s) and options.

    method 'omw' -> decon_lucy_omw_function(chunk, PSF_f, PSF_b, NUMIT, ...);
    method 'rl'  -> decon_lucy_function(chunk, PSF_f, NUMIT, ...) (PSF_b ignored).
    Returns the list of deconvolved chunks (numpy for host chunks, CUDA
    tensors for device chunks)."""
    o = _parse(kwargs)
    if not o["useGPU"]:
        raise ValueError("useGPU=false: this engine implements the GPU branch only")
    m = {"omw": 1, "rl": 0, "simplified": 0}[method]
    opts, _ = _opts_struct(o, (1, 1, 1), False)
    if opts.has_bbox:
        raise ValueError("bbox is per chunk: crop the outputs")
    ctx = o["ctx"] or _lib.default_context()
    keep, items, outs = [], (PkBatchItem * len(chunks))(), []
    for i, c in enumerate(chunks):
        obj, vin, kind = _vol_of(c)
        d = tuple(vin.dims)
        if kind == "torch":
            import torch
            t = torch.empty((d[2], d[1], d[0]), dtype=torch.uint16 if o["save16bit"] else torch.float32,
                            device=obj.device).permute(2, 1, 0)
            vout = PkVol(ctypes.c_void_p(t.data_ptr()), PK_UINT16 if o["save16bit"] else PK_SINGLE, PK_MEM_DEVICE,
                         (ctypes.c_int64 * 3)(*d))
            outs.append(t)
        else:
            a = np.empty(d, dtype=np.uint16 if o["save16bit"] else np.float32, order="F")
            vout = PkVol(ctypes.c_void_p(a.ctypes.data), PK_UINT16 if o["save16bit"] else PK_SINGLE, PK_MEM_HOST,
                         (ctypes.c_int64 * 3)(*d))
            outs.append(a)
        keep.append(obj)
        items[i].inp = vin
        items[i].out = vout
    if any(_is_torch(c) for c in chunks):
        import torch
        torch.cuda.synchronize()
    pf, fd = _psf_array(PSF_f)
    pb, bd = _psf_array(PSF_b if m == 1 else PSF_f)
    ctx.check(_lib.load().pk_decon_batch(ctx.handle, m, items, len(chunks), ctypes.c_void_p(pf.ctypes.data), i64x3(fd),
                                         ctypes.c_void_p(pb.ctypes.data), i64x3(bd), int(NUMIT), ctypes.byref(opts)))
    return outs

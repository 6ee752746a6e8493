"""Large-volume chunking: the chunk-coordinate generator of the zarr / large
volume decon drivers and an in-memory driver that deconvolves every chunk in
one batched engine call (pk_decon_batch) and writes each chunk's region back.

Bounding boxes follow MATLAB: 1-based, inclusive, [y0 x0 z0 y1 x1 z1] rows
(dim 1 = y, contiguous)."""
from __future__ import annotations

import numpy as np

from .batch import decon_batch
from .decon import _is_torch

__all__ = ["XR_zarrChunkCoordinatesExtraction", "decon_large_in_memory", "matlab_round"]


def matlab_round(x):
    """MATLAB round: half away from zero."""
    a = np.asarray(x, dtype=np.float64)
    return (np.sign(a) * np.floor(np.abs(a) + 0.5)).astype(np.int64)


def _vec3(v, name):
    a = np.atleast_1d(np.asarray(v, dtype=np.int64))
    if a.size == 1:
        a = np.repeat(a, 3)
    if a.size != 3:
        raise ValueError(f"{name} must have 1 or 3 elements")
    return a


def XR_zarrChunkCoordinatesExtraction(volSize, batchSize=(1024, 1024, 1024), BlockSize=(256, 256, 256), bbox=None,
                                      sameBatchSize=True, BorderSize=50):
    """[batchBBoxes, regionBBoxes, localBBoxes] for a volume split into
    batches of whole blocks (stitch/XR_zarrChunkCoordinatesExtraction.m:1-100).

    The processed world is bbox (or the whole volume).  batchSize and
    BlockSize are clipped to it and batchSize rounded up to whole blocks
    (:47-51); regions tile the world in column-major batch order (:53-63);
    with sameBatchSize each batch starts batchSize - 1 before its region's
    end (:68-70); batches grow by BorderSize, clipped to the volume
    (:72-73); localBBoxes place each region inside its batch (:76-78);
    regions are returned relative to the world start (:81).  Arrays are
    int64 (n, 6)."""
    vol = _vec3(volSize, "volSize")
    if vol.size != 3 or np.any(vol < 1):
        raise ValueError(f"The volume size {list(vol)} is not valid for a 3d volume.")
    bs = _vec3(batchSize, "batchSize")
    blk = _vec3(BlockSize, "BlockSize")
    border = _vec3(BorderSize, "BorderSize")
    if np.any(bs < 1):
        raise ValueError("batchSize must be >= 1")
    if bbox is None or len(bbox) == 0:
        wd_start = np.ones(3, dtype=np.int64)
        wd_end = vol.copy()
    else:
        b = np.asarray(bbox, dtype=np.int64).reshape(6)
        wd_start, wd_end = b[:3].copy(), b[3:].copy()
    wd_size = wd_end - wd_start + 1
    bs = np.minimum(wd_size, bs)
    blk = np.minimum(wd_size, blk)
    bs = -(-bs // blk) * blk  # whole blocks per batch
    nsub = -(-wd_size // bs)
    nb = int(np.prod(nsub))
    # batch subscripts in MATLAB ndgrid / column-major order (dim 1 fastest)
    k = np.arange(nb, dtype=np.int64)
    subs = np.stack([k % nsub[0], (k // nsub[0]) % nsub[1], k // (nsub[0] * nsub[1])], axis=1)  # 0-based
    region = np.zeros((nb, 6), dtype=np.int64)
    region[:, :3] = subs * bs + wd_start
    region[:, 3:] = np.minimum(region[:, :3] + bs - 1, wd_end)
    batch = region.copy()
    if sameBatchSize:
        batch[:, :3] = batch[:, 3:] - bs + 1
    batch[:, :3] = np.maximum(1, batch[:, :3] - border)
    batch[:, 3:] = np.minimum(vol, batch[:, 3:] + border)
    local = np.zeros((nb, 6), dtype=np.int64)
    local[:, :3] = region[:, :3] - batch[:, :3] + 1
    local[:, 3:] = local[:, :3] + (region[:, 3:] - region[:, :3])
    region = region - np.concatenate([wd_start, wd_start]) + 1
    return batch, region, local


def _crop(x, b):
    """Volume x cropped to the 1-based inclusive box b (crop3d_mex.c:54-59
    convention), column-major: a numpy array or a CUDA tensor."""
    s = (slice(int(b[0]) - 1, int(b[3])), slice(int(b[1]) - 1, int(b[4])), slice(int(b[2]) - 1, int(b[5])))
    if _is_torch(x):
        c = x[s]
        return c.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    return np.asfortranarray(x[s])


def edge_erosion_mask(im, EdgeErosion, ctx=None):
    """The whole-volume edge-erosion mask of RLdecon_large_in_memory.m:150-155:
    bw = im > 0 with the six faces cleared, kept where bwdist(~bw) >
    EdgeErosion - 1 (pk_edge_erosion_mask on the device).  Boolean numpy
    array for a host volume, bool CUDA tensor for a device volume."""
    import ctypes

    from . import _lib
    from ._lib import PK_MEM_DEVICE, PK_MEM_HOST, PK_SINGLE, PK_UINT8, PK_UINT16, PkVol
    shape = tuple(int(n) for n in (im.shape if _is_torch(im) else np.shape(im)))
    if len(shape) != 3:
        raise ValueError("a 3-D volume is required")
    ctx = ctx or _lib.default_context()
    if _is_torch(im) and im.is_cuda:
        import torch
        t = im if im.dtype in (torch.uint16, torch.float32) else im.to(torch.float32)
        if tuple(t.stride()) != (1, shape[0], shape[0] * shape[1]):
            t = t.permute(2, 1, 0).contiguous().permute(2, 1, 0)
        m = torch.empty((shape[2], shape[1], shape[0]), dtype=torch.uint8, device=t.device).permute(2, 1, 0)
        vin = PkVol(ctypes.c_void_p(t.data_ptr()), PK_UINT16 if t.dtype == torch.uint16 else PK_SINGLE,
                    PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*shape))
        vm = PkVol(ctypes.c_void_p(m.data_ptr()), PK_UINT8, PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*shape))
        torch.cuda.current_stream(t.device).synchronize()
        ctx.check(_lib.load().pk_edge_erosion_mask(ctx.handle, ctypes.byref(vin), int(EdgeErosion), ctypes.byref(vm)))
        return m.bool()
    a = np.asarray(im.numpy() if _is_torch(im) else im)
    if a.dtype not in (np.uint16, np.float32):
        a = a.astype(np.float32)
    a = np.asfortranarray(a)
    m = np.empty(shape, dtype=np.uint8, order="F")
    vin = PkVol(ctypes.c_void_p(a.ctypes.data), PK_UINT16 if a.dtype == np.uint16 else PK_SINGLE, PK_MEM_HOST,
                (ctypes.c_int64 * 3)(*shape))
    vm = PkVol(ctypes.c_void_p(m.ctypes.data), PK_UINT8, PK_MEM_HOST, (ctypes.c_int64 * 3)(*shape))
    ctx.check(_lib.load().pk_edge_erosion_mask(ctx.handle, ctypes.byref(vin), int(EdgeErosion), ctypes.byref(vm)))
    return m.astype(bool)


def decon_large_in_memory(im, PSF_f, PSF_b, NUMIT, method="omw", batchSize=(512, 512, 512), EdgeErosion=8,
                          ErodeMaskfile=None, **kwargs):
    """Deconvolve a volume larger than one chunk, as the in-memory large-volume
    driver does (decon/RLdecon_large_in_memory.m:171-262): batches of
    batchSize (BlockSize = batchSize, sameBatchSize false) with a border of
    round((size(psf) + 10) / 2) on every side (:171-175), each deconvolved and
    its region written into the output (:189-256).  With EdgeErosion > 0 (the
    reference's default 8) the assembled volume is multiplied by
    edge_erosion_mask(im, EdgeErosion) (:141-162, 271-274); an existing
    ErodeMaskfile (TIFF) replaces that mask (:78-80, 276-280).

    Here every batch goes through ONE pk_decon_batch call (one launch per pass
    and block shape over all batches, the PSF centre-cropped for batches
    smaller than it, RLdecon.m:264-270) instead of one decon call per batch.
    kwargs are decon_batch's (Background, save16bit, dampFactor, ...; bbox is
    not accepted: the output is the whole volume).  Returns the volume
    (numpy for a host input, a column-major CUDA tensor for a device input)."""
    import os
    if "bbox" in {k.lower() for k in kwargs}:
        raise ValueError("bbox: the large-volume driver writes the whole volume")
    erode_file = ErodeMaskfile if ErodeMaskfile and os.path.exists(str(ErodeMaskfile)) else None
    if erode_file:
        EdgeErosion = 0
    mask = edge_erosion_mask(im, EdgeErosion, kwargs.get("ctx")) if EdgeErosion and EdgeErosion > 0 else None
    shape = tuple(int(n) for n in (im.shape if _is_torch(im) else np.shape(im)))
    if len(shape) != 3:
        raise ValueError("a 3-D volume is required")
    psf_shape = np.shape(PSF_f.detach().cpu().numpy() if _is_torch(PSF_f) else PSF_f)
    border = matlab_round((np.asarray(psf_shape, dtype=np.float64) + 10) / 2)
    bs = _vec3(batchSize, "batchSize")
    batch, region, _ = XR_zarrChunkCoordinatesExtraction(shape, batchSize=bs, BlockSize=bs, sameBatchSize=False,
                                                         BorderSize=border)
    chunks = [_crop(im, b) for b in batch]
    outs = decon_batch(chunks, PSF_f, PSF_b, NUMIT, method=method, **kwargs)
    # region of each output inside its batch: baStart = obStart - ibStart + 1 (:228-230)
    if _is_torch(im):
        import torch
        dt = outs[0].dtype
        out = torch.empty((shape[2], shape[1], shape[0]), dtype=dt, device=im.device).permute(2, 1, 0)
    else:
        out = np.empty(shape, dtype=outs[0].dtype, order="F")
    for b, r, o in zip(batch, region, outs):
        lo = r[:3] - b[:3]  # 0-based start of the region inside the batch
        hi = lo + (r[3:] - r[:3]) + 1
        out[r[0] - 1:r[3], r[1] - 1:r[4], r[2] - 1:r[5]] = o[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    if erode_file:
        from .tiff_io import readtiff
        m = np.asarray(readtiff(erode_file))
        if m.shape != shape:
            raise ValueError(f"ErodeMaskfile size {m.shape} differs from the volume's {shape}")
        mask = m != 0
        if _is_torch(out):
            import torch
            mask = torch.from_numpy(np.ascontiguousarray(mask.transpose(2, 1, 0))).to(out.device).permute(2, 1, 0)
    if mask is not None:  # im .* cast(im_bw_erode, class(im))
        if _is_torch(out):
            import torch
            tgt = out.view(torch.int16) if out.dtype == torch.uint16 else out  # same bits; uint16 lacks masked_fill
            tgt.masked_fill_(~mask, 0)
        else:
            out[~mask] = 0
    return out

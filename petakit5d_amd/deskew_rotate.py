"""deskewRotateFrame3D (microscopeDataProcessing/deskew_rotate/deskewRotateFrame3D.m:1-260):
deskew and rotation of a stage-scan volume in one resampling, the reference's
default after decon (RLdecon.m:22 DSRCombined = true, RLdecon.m:489-495).
Runs on the device through libpkdecon.so's pk_deskew_rotate_frame3d:
skewed-space interpolation to a finer slice step when the slice-to-slice
shift exceeds xStepThresh, then the row-preserving bilinear warp of the
reference's fast path (volume_deskew_rotate_warp_mex.cpp); objective scans
take imwarp's trilinear form.  interpMethod 'cubic' leaves the fast path as
the reference does: the skewed-space interpolation when the step is large,
then imwarp's cubic convolution of the whole transform
(pk_deskew_rotate_frame3d_cubic).  A resampleFactor other than ones takes the
same imwarp form with the resampling folded into the transform
(pk_deskew_rotate_frame3d_resample), linear or cubic, and so does gpuProcess
(the reference runs imwarp on a gpuArray then, deskewRotateFrame3D.m:73)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PK_MEM_DEVICE, PK_SINGLE, PK_UINT16, PkVol, i64x3

__all__ = ["deskewRotateFrame3D", "deskew_rotate_out_size"]


def _resample(resampleFactor):
    """resampleFactor as 3 doubles (x y z), or None when it is empty or all
    ones (deskewRotateFrame3D.m:72-73: the fast path then applies)."""
    if resampleFactor is None or not np.size(resampleFactor):
        return None
    rs = np.asarray(resampleFactor, dtype=np.float64).reshape(-1)
    if rs.size != 3:
        raise ValueError("deskewRotateFrame3D: resampleFactor must have 3 elements (x y z)")
    if not np.all(np.isfinite(rs)) or np.any(rs <= 0):
        raise ValueError("deskewRotateFrame3D: resampleFactor must be positive")
    if np.all(rs == 1):
        return None
    return (ctypes.c_double * 3)(*rs.tolist())


def deskew_rotate_out_size(shape, angle, dz, xyPixelSize, reverse=False, objectiveScan=False, xStepThresh=2.0,
                           resampleFactor=None):
    """outSize [ny nx' nz'] of deskewRotateFrame3D (deskewRotateFrame3D.m:61-69;
    resampled: round(outSize ./ resampleFactor([2 1 3])), :189)."""
    d = tuple(int(n) for n in shape) + (1,) * (3 - len(shape))
    out = (ctypes.c_int64 * 3)()
    rs = _resample(resampleFactor)
    if rs is not None:
        rc = _lib.load().pk_deskew_rotate_resample_out_dims(
            i64x3(d), ctypes.c_double(angle), ctypes.c_double(dz), ctypes.c_double(xyPixelSize), int(bool(reverse)),
            int(bool(objectiveScan)), ctypes.c_double(xStepThresh), rs, out)
    else:
        rc = _lib.load().pk_deskew_rotate_out_dims(i64x3(d), ctypes.c_double(angle), ctypes.c_double(dz),
                                                   ctypes.c_double(xyPixelSize), int(bool(reverse)),
                                                   int(bool(objectiveScan)), ctypes.c_double(xStepThresh), out)
    if rc != 0:
        raise _lib.PkDeconError(rc, "deskewRotateFrame3D: invalid size or geometry")
    return tuple(int(v) for v in out)


def deskewRotateFrame3D(vol, angle, dz, xyPixelSize, reverse=False, Crop=True, bbox=None, objectiveScan=False,
                        xStepThresh=2.0, resampleFactor=None, gpuProcess=False, save16bit=False,
                        interpMethod="linear", ctx=None):
    """volout = deskewRotateFrame3D(vol, angle, dz, xyPixelSize, 'reverse', reverse,
    'Crop', Crop, 'bbox', bbox, 'objectiveScan', objectiveScan, 'xStepThresh',
    xStepThresh, 'resampleFactor', resampleFactor, 'save16bit', save16bit,
    'interpMethod', interpMethod).

    vol: uint16 or single, numpy (column-major result) or a CUDA tensor (CUDA
    result).  The output is single, or uint16 for a uint16 input with
    save16bit (the reference's mex does the same).  'Crop' is accepted and,
    as in the reference, has no effect.  gpuProcess leaves the fast path as
    the reference does (:73): skewed-space interpolation when the step is
    large, then imwarp's trilinear form (:210).  resampleFactor (x y z; [] or
    ones: none) other than ones takes the same imwarp form with the
    resampling folded in (:183-210).  With 'cubic', gpuProcess or a
    resampling factor the output has imwarp's class of what it resamples
    (:98-110, 210):
    uint16 for a uint16 volume that is not interpolated in skewed space, or
    is with save16bit; single otherwise."""
    import math
    import torch
    interp = str(interpMethod).lower()
    if interp not in ("linear", "cubic"):
        raise ValueError("deskewRotateFrame3D: interpMethod must be 'linear' or 'cubic'")
    rs = _resample(resampleFactor)
    on_dev = isinstance(vol, torch.Tensor) and vol.is_cuda
    if on_dev:
        t = vol
    else:
        a = np.asarray(vol.detach().cpu().numpy() if isinstance(vol, torch.Tensor) else vol)
        if a.dtype not in (np.uint16, np.float32):
            a = a.astype(np.float32)
        a = np.asfortranarray(a.reshape(a.shape + (1,) * (3 - a.ndim)))
        t = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    if t.dtype not in (torch.uint16, torch.float32):
        t = t.to(torch.float32)
    d = tuple(int(n) for n in t.shape)
    if t.stride() != (1, d[0], d[0] * d[1]):
        t = t.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    out_size = deskew_rotate_out_size(d, angle, dz, xyPixelSize, reverse, objectiveScan, xStepThresh, resampleFactor)
    box = None
    od = out_size
    if bbox is not None and np.size(bbox):
        b = [int(v) for v in np.asarray(bbox).reshape(-1)]
        if len(b) != 6:
            raise ValueError("bbox must be [s1 s2 s3 e1 e2 e3]")
        box = (ctypes.c_int64 * 6)(*b)
        od = (b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1)
    u16 = t.dtype == torch.uint16 and bool(save16bit)
    imwarp_form = interp == "cubic" or rs is not None or bool(gpuProcess)
    if imwarp_form and t.dtype == torch.uint16:
        do_interp = not objectiveScan and abs(math.cos(angle * math.pi / 180) * dz / xyPixelSize) > xStepThresh
        u16 = bool(save16bit) or not do_interp
    out = torch.empty((od[2], od[1], od[0]), dtype=torch.uint16 if u16 else torch.float32,
                      device=t.device).permute(2, 1, 0)
    vin = PkVol(ctypes.c_void_p(t.data_ptr()), PK_UINT16 if t.dtype == torch.uint16 else PK_SINGLE, PK_MEM_DEVICE,
                (ctypes.c_int64 * 3)(*d))
    vout = PkVol(ctypes.c_void_p(out.data_ptr()), PK_UINT16 if u16 else PK_SINGLE, PK_MEM_DEVICE,
                 (ctypes.c_int64 * 3)(*od))
    ctx = ctx or _lib.default_context()
    torch.cuda.current_stream(t.device).synchronize()
    L = _lib.load()
    args = (ctypes.c_double(angle), ctypes.c_double(dz), ctypes.c_double(xyPixelSize), int(bool(reverse)),
            int(bool(objectiveScan)), ctypes.c_double(xStepThresh))
    if rs is not None or gpuProcess:
        flag = (1 if interp == "cubic" else 0) + (2 if gpuProcess else 0)
        ctx.check(L.pk_deskew_rotate_frame3d_resample(ctx.handle, ctypes.byref(vin), *args, rs, flag, box,
                                                      ctypes.byref(vout)))
    else:
        fn = L.pk_deskew_rotate_frame3d_cubic if interp == "cubic" else L.pk_deskew_rotate_frame3d
        ctx.check(fn(ctx.handle, ctypes.byref(vin), *args, box, ctypes.byref(vout)))
    if on_dev:
        return out
    return np.asfortranarray(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0))

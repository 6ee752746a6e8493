"""deskewFrame3D for the deconvolved volume (pointDetection/llsmtools/
transforms/deskewFrame3D.m:26-70, the step after decon at
microscopeDataProcessing/decon/RLdecon.m:448-449), on the device through
libpkdecon.so's pk_deskew_frame3d (linear interpolation, fill value 0), or
with 'Interp', 'cubic' through pk_affine_warp3d_cubic (imwarp's cubic
convolution of the same shear)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import i64x3

__all__ = ["deskewFrame3D", "deskew_out_size", "deskew_transform"]


def deskew_out_size(shape, angle, dz, xyPixelSize, reverse=False, Crop=False):
    """[ny, nxOut, nz] of deskewFrame3D for a volume of MATLAB size shape."""
    d = [int(n) for n in shape[:3]] + [1] * (3 - len(shape[:3]))
    out = (ctypes.c_int64 * 3)()
    rc = _lib.load().pk_deskew_out_dims(i64x3(d), ctypes.c_double(angle), ctypes.c_double(dz),
                                        ctypes.c_double(xyPixelSize), int(bool(reverse)), int(bool(Crop)), out)
    if rc != 0:
        raise _lib.PkDeconError(rc, "deskew: invalid size or geometry")
    return tuple(int(v) for v in out)


def deskew_transform(shape, angle, dz, xyPixelSize, reverse=False, Crop=False):
    """(S, outSize) of deskewFrame3D (deskewFrame3D.m:40-69): S the 4x4 shear
    in imwarp's row-vector convention, outSize [ny nxOut nz]."""
    import math
    ny, nx, nz = (int(n) for n in shape)
    theta = angle * math.pi / 180
    dx = math.cos(theta) * dz / xyPixelSize
    if not reverse:
        xshift, xstep = -dx, dx
    else:
        xshift, xstep = dx + math.ceil((nz - 1) * dx), -dx
    nx_out = math.ceil((nz - 1) * dx) + nx
    if Crop:
        crop = nx // 2 if nx_out >= 2 * nx else math.floor(math.ceil((nz - 1) * dx) / 2)
        xshift -= crop
        nx_out -= 2 * crop
    S = np.eye(4)
    S[2, 0] = xstep
    S[3, 0] = xshift
    return S, (ny, int(nx_out), nz)


def deskewFrame3D(vol, angle, dz, xyPixelSize, reverse=False, Crop=False, Interp="linear", ctx=None):
    """volout = deskewFrame3D(vol, angle, dz, xyPixelSize, reverse, 'Crop', Crop, 'Interp', Interp).

    vol: numpy array (MATLAB size order, any real type; converted to single
    as RLdecon.m:448 does) or a float32 CUDA tensor; the result has the same
    kind (numpy column-major, or a column-major CUDA tensor).  Interp
    'linear' or 'cubic' (deskewFrame3D.m:36)."""
    import torch
    interp = str(Interp).lower()
    if interp not in ("linear", "cubic"):
        raise ValueError("deskewFrame3D: Interp must be 'linear' or 'cubic'")
    on_dev = isinstance(vol, torch.Tensor) and vol.is_cuda
    if on_dev:
        t = vol.to(torch.float32)
    else:
        a = np.asarray(vol.detach().cpu().numpy() if isinstance(vol, torch.Tensor) else vol)
        a = np.asfortranarray(a, dtype=np.float32)
        t = torch.from_numpy(a.reshape(a.shape + (1,) * (3 - a.ndim)).transpose(2, 1, 0).copy()).cuda()
        t = t.permute(2, 1, 0)
    if t.dim() == 2:
        t = t.unsqueeze(2)
    d = tuple(int(n) for n in t.shape)
    if t.stride() != (1, d[0], d[0] * d[1]):
        t = t.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    od = deskew_out_size(d, angle, dz, xyPixelSize, reverse, Crop)
    out = torch.empty((od[2], od[1], od[0]), dtype=torch.float32, device=t.device).permute(2, 1, 0)
    ctx = ctx or _lib.default_context()
    torch.cuda.current_stream(t.device).synchronize()
    if interp == "cubic":
        S, od2 = deskew_transform(d, angle, dz, xyPixelSize, reverse, Crop)
        assert tuple(od2) == tuple(od)
        inv = np.ascontiguousarray(np.linalg.inv(S), dtype=np.float64)
        ctx.check(_lib.load().pk_affine_warp3d_cubic(ctx.handle, ctypes.c_void_p(t.data_ptr()), i64x3(d),
                                                     inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), i64x3(od),
                                                     ctypes.c_void_p(out.data_ptr())))
        if on_dev:
            return out
        return np.asfortranarray(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0))
    ctx.check(_lib.load().pk_deskew_frame3d(ctx.handle, ctypes.c_void_p(t.data_ptr()), i64x3(d),
                                            ctypes.c_double(angle), ctypes.c_double(dz),
                                            ctypes.c_double(xyPixelSize), int(bool(reverse)), int(bool(Crop)),
                                            ctypes.c_void_p(out.data_ptr())))
    if on_dev:
        return out
    return np.asfortranarray(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0))

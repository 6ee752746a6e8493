"""deskewFrame3D for the deconvolved volume (pointDetection/llsmtools/
transforms/deskewFrame3D.m:26-70, the step after decon at
microscopeDataProcessing/decon/RLdecon.m:448-449), on the device through
libpkdecon.so's pk_deskew_frame3d.  Linear interpolation, fill value 0;
'cubic' is not served."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import i64x3

__all__ = ["deskewFrame3D", "deskew_out_size"]


def deskew_out_size(shape, angle, dz, xyPixelSize, reverse=False, Crop=False):
    """[ny, nxOut, nz] of deskewFrame3D for a volume of MATLAB size shape."""
    d = [int(n) for n in shape[:3]] + [1] * (3 - len(shape[:3]))
    out = (ctypes.c_int64 * 3)()
    rc = _lib.load().pk_deskew_out_dims(i64x3(d), ctypes.c_double(angle), ctypes.c_double(dz),
                                        ctypes.c_double(xyPixelSize), int(bool(reverse)), int(bool(Crop)), out)
    if rc != 0:
        raise _lib.PkDeconError(rc, "deskew: invalid size or geometry")
    return tuple(int(v) for v in out)


def deskewFrame3D(vol, angle, dz, xyPixelSize, reverse=False, Crop=False, Interp="linear", ctx=None):
    """volout = deskewFrame3D(vol, angle, dz, xyPixelSize, reverse, 'Crop', Crop, 'Interp', Interp).

    vol: numpy array (MATLAB size order, any real type; converted to single
    as RLdecon.m:448 does) or a float32 CUDA tensor; the result has the same
    kind (numpy column-major, or a column-major CUDA tensor)."""
    import torch
    if str(Interp).lower() != "linear":
        raise NotImplementedError("deskewFrame3D: only 'linear' interpolation is served")
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
    ctx.check(_lib.load().pk_deskew_frame3d(ctx.handle, ctypes.c_void_p(t.data_ptr()), i64x3(d),
                                            ctypes.c_double(angle), ctypes.c_double(dz),
                                            ctypes.c_double(xyPixelSize), int(bool(reverse)), int(bool(Crop)),
                                            ctypes.c_void_p(out.data_ptr())))
    if on_dev:
        return out
    return np.asfortranarray(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0))

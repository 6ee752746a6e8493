"""rotateFrame3D (pointDetection/llsmtools/transforms/rotateFrame3D.m:26-100):
rotation of a deskewed (or objective-scan) volume about the y axis with the z
anisotropy zxRatio, resampled on the device by libpkdecon.so's
pk_affine_warp3d (imwarp 'linear', fill value 0) or pk_affine_warp3d_cubic
(imwarp 'cubic').  The transform, the output size and the automatic crop
follow the reference line by line."""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import i64x3

__all__ = ["rotateFrame3D", "rotate_transform"]


def _mround(x):
    return math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)


def _translate(t):
    m = np.eye(4)
    m[3, :3] = t
    return m


def rotate_transform(shape, angle, zxRatio, reverse=False, Crop=True, resample=None, ObjectiveScan=False,
                     outSize=None, proj=None):
    """(T, outSize) of rotateFrame3D for a volume of MATLAB size shape: T the
    4x4 forward transform in imwarp's row-vector convention ([x y z 1] * T),
    outSize [ny' nx' nz'].  proj: the volume's max over dim 1 as an nx x nz
    array (needed for the crop decision of stage scans)."""
    ny, nx, nz = (int(n) for n in shape)
    theta = angle * math.pi / 180
    if reverse:
        theta = -theta
    do_crop, w, h = False, None, None
    if not ObjectiveScan:  # rotateFrame3D.m:46-66
        pz = (np.asarray(proj) != 0).T.astype(np.int8)  # squeeze(max(vol, [], 1))' ~= 0: nz x nx
        d = np.diff(pz, axis=1)  # diff(proj, 1, 2)
        xs, zs = np.nonzero(d.T)  # [z, x] = find(proj ~= 0), in find's column-major order
        z, x = zs + 1.0, xs + 1.0
        a, b = math.sin(-theta), math.cos(-theta)
        if z.size:
            c = a * x + b * z * zxRatio
            h = _mround(c.max() - c.min()) - 4
            do_crop = h < max(abs(a * (nx - 1) + b * (nz - 1) * zxRatio), abs(a * (nx - 1) + b * (1 - nz) * zxRatio))
            cx = b * x + (-a) * z * zxRatio
            w = _mround(cx.max() - cx.min()) - 4
    center = np.array([ny, nx, nz], dtype=np.float64) + 1
    center /= 2
    T1 = _translate(-center[[1, 0, 2]])
    S = np.diag([1.0, 1.0, float(zxRatio), 1.0])
    ct, st = math.cos(theta), math.sin(theta)
    R = np.array([[ct, 0, -st, 0], [0, 1, 0, 0], [st, 0, ct, 0], [0, 0, 0, 1]], dtype=np.float64)
    if outSize is None:
        if Crop and do_crop:
            out = [ny, w, h]
        else:
            out = [ny, nx * ct + nz * zxRatio / math.sin(abs(theta)), nz * zxRatio * ct + nx * math.sin(abs(theta))]
        out = [_mround(v) for v in out]
    else:
        out = [int(v) for v in outSize]
    out = np.array(out, dtype=np.float64)
    T2 = _translate((out[[1, 0, 2]] + 1) / 2)
    if resample is not None and len(resample):
        rs = np.asarray(resample, dtype=np.float64).reshape(-1)
        rs = np.repeat(rs, 3) if rs.size == 1 else rs
        RT1 = _translate(-(out[[1, 0, 2]] + 1) / 2)
        RS = np.diag([1 / rs[0], 1 / rs[1], 1 / rs[2], 1.0])
        out = np.array([_mround(v) for v in out / rs[[1, 0, 2]]], dtype=np.float64)
        RT2 = _translate((out[[1, 0, 2]] + 1) / 2)
    else:
        RT1 = RS = RT2 = np.eye(4)
    T = (T1 @ S @ R @ T2) @ (RT1 @ RS @ RT2)
    return T, tuple(int(v) for v in out)


def rotateFrame3D(vol, angle, zxRatio, reverse=False, Crop=True, resample=None, ObjectiveScan=False, outSize=None,
                  Interp="linear", ctx=None):
    """volout = rotateFrame3D(vol, angle, zxRatio, reverse, 'Crop', Crop,
    'resample', resample, 'ObjectiveScan', ObjectiveScan, 'outSize', outSize,
    'Interp', Interp): numpy in -> column-major numpy out, CUDA tensor in ->
    CUDA tensor out.  Interp 'linear' or 'cubic' (rotateFrame3D.m:40)."""
    import torch
    interp = str(Interp).lower()
    if interp not in ("linear", "cubic"):
        raise ValueError("rotateFrame3D: Interp must be 'linear' or 'cubic'")
    on_dev = isinstance(vol, torch.Tensor) and vol.is_cuda
    if on_dev:
        t = vol.to(torch.float32)
    else:
        a = np.asarray(vol.detach().cpu().numpy() if isinstance(vol, torch.Tensor) else vol)
        a = np.asfortranarray(a, dtype=np.float32)
        t = torch.from_numpy(a.reshape(a.shape + (1,) * (3 - a.ndim)).transpose(2, 1, 0).copy()).cuda()
        t = t.permute(2, 1, 0)
    d = tuple(int(n) for n in t.shape)
    if t.stride() != (1, d[0], d[0] * d[1]):
        t = t.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    proj = None if ObjectiveScan else t.amax(dim=0).cpu().numpy()
    T, od = rotate_transform(d, angle, zxRatio, reverse, Crop, resample, ObjectiveScan, outSize, proj)
    if min(od) < 1:
        raise ValueError(f"rotateFrame3D: output size {od}")
    inv = np.ascontiguousarray(np.linalg.inv(T), dtype=np.float64)
    out = torch.empty((od[2], od[1], od[0]), dtype=torch.float32, device=t.device).permute(2, 1, 0)
    ctx = ctx or _lib.default_context()
    torch.cuda.current_stream(t.device).synchronize()
    fn = _lib.load().pk_affine_warp3d_cubic if interp == "cubic" else _lib.load().pk_affine_warp3d
    ctx.check(fn(ctx.handle, ctypes.c_void_p(t.data_ptr()), i64x3(d), inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                 i64x3(od), ctypes.c_void_p(out.data_ptr())))
    if on_dev:
        return out
    return np.asfortranarray(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0))

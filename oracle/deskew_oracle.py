"""numpy restatement of deskewFrame3D (linear interpolation).

TEST INFRASTRUCTURE ONLY: the checker for the engine's deskew.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.

Follows pointDetection/llsmtools/transforms/deskewFrame3D.m:40-70: dx =
cos(angle pi / 180) dz / xyPixelSize; (xstep, xshift) = (dx, -dx), or (-dx,
dx + ceil((nz - 1) dx)) when reversed; nxOut = ceil((nz - 1) dx) + nx; 'Crop'
trims floor(nx / 2) or floor(ceil((nz - 1) dx) / 2) columns per side.  The
imwarp resampling (OutputView imref3d([ny nxOut nz], 1, 1, 1), FillValues 0)
is restated as: output column c of plane p reads input column
x = c - xstep p - xshift (1-based), linear interpolation over the input padded
with one zero column on each side, 0 outside [0, nx + 1].  The padding rule is
imwarp's edge blend as documented for linear resampling; it is pinned here by
closed forms (integer shifts, linear ramps), not by MATLAB output."""
import math

import numpy as np


def geometry(shape, angle, dz, xy, reverse=False, crop=False):
    ny, nx, nz = shape
    dx = math.cos(angle * math.pi / 180) * dz / xy
    span = math.ceil((nz - 1) * dx)
    xstep, xshift = (-dx, dx + span) if reverse else (dx, -dx)
    nx_out = span + nx
    if crop:
        cw = nx // 2 if nx_out >= 2 * nx else math.floor(span / 2)
        xshift -= cw
        nx_out -= 2 * cw
    return nx_out, xstep, xshift


def deskew(vol, angle, dz, xy, reverse=False, crop=False):
    """Float64 result of deskewFrame3D(single(vol), ...)."""
    v = np.asarray(vol, dtype=np.float32).astype(np.float64)
    ny, nx, nz = v.shape
    nx_out, xstep, xshift = geometry(v.shape, angle, dz, xy, reverse, crop)
    pad = np.zeros((ny, nx + 2, nz))
    pad[:, 1:nx + 1, :] = v  # padded column k holds input column k (1-based), 0 and nx + 1 are zero
    out = np.zeros((ny, nx_out, nz))
    c = np.arange(1, nx_out + 1, dtype=np.float64)
    for p in range(1, nz + 1):
        x = c - xstep * p - xshift
        ok = (x >= 0) & (x <= nx + 1)
        x0 = np.floor(x[ok]).astype(np.int64)
        f = x[ok] - x0
        a = pad[:, x0, p - 1]
        b = pad[:, np.minimum(x0 + 1, nx + 1), p - 1] * (x0 + 1 <= nx + 1)
        out[:, np.nonzero(ok)[0], p - 1] = a + f * (b - a)
    return out

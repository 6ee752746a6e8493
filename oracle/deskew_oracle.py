"""numpy restatement of deskewFrame3D (linear interpolation), of
deskewRotateFrame3D (linear fast path; 'cubic' and resampleFactor through
imwarp's form, deskew_rotate_general), and of imwarp's cubic resampling
(cubic_warp).

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


# ---------------------------------------------------------------------------------------------------------------------
# deskewRotateFrame3D (microscopeDataProcessing/deskew_rotate/deskewRotateFrame3D.m:50-253), linear interpolation, no
# resampling: the reference's fast path -- skewed-space interpolation to a finer z step when the slice-to-slice
# shift exceeds xStepThresh (skewed_space_interp_defined_stepsize.m:24-100), then the row-preserving bilinear warp of
# volume_deskew_rotate_warp_mex.cpp:217-282 with the backward matrix of deskewRotateFrame3D.m:212-219.
def _mround(x):
    return math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)


def dsr_geometry(shape, angle, dz, xy, reverse=False, objectiveScan=False, xStepThresh=2.0, resample=None):
    """(tmat, outSize, xstep, stepsize, T) -- tmat the 4x4 backward matrix over
    0-based (dim1, dim2, dim3) coordinates (the fast path's), outSize [ny nx'
    nz'], xstep the original |dx|, stepsize the interpolated slice step (1:
    none), T imwarp's forward matrix.  resample (x y z) other than ones
    (deskewRotateFrame3D.m:183-196): T gains RT1 * RS * RT2 -- the output box
    centre to the origin, RS = diag(1 ./ resample), the origin to the centre of
    the box round(outSize ./ resample([2 1 3])), which becomes outSize."""
    ny, nx, nz = (int(n) for n in shape)
    theta = angle * math.pi / 180
    dx = math.cos(theta) * dz / xy
    zAniso = dz / xy if objectiveScan else math.sin(abs(theta)) * dz / xy
    if not objectiveScan:
        out = [ny, (nx - 1) * math.cos(theta) + (nz - 1) * zAniso / math.sin(abs(theta)),
               (nx - 1) * math.sin(abs(theta)) - 4]
    else:
        out = [ny, nx * math.cos(theta) + nz * zAniso * math.sin(abs(theta)),
               nz * zAniso * math.cos(theta) + nx * math.sin(abs(theta))]
    out = [_mround(v) for v in out]
    stepsize, xstep0 = 1.0, abs(dx)
    if not objectiveScan and abs(dx) > xStepThresh:
        if abs(dx) / xStepThresh < 1.5:
            thr = xy * xStepThresh / math.cos(theta)
            dzout = dz / math.ceil(dz / thr)
            counter = 1
            while dzout / thr < 0.4:
                dzout = dz / (math.ceil((dz / thr) / (1 / (counter + 1))) * (1 / (counter + 1)))
                counter += 1
        else:
            dzout = dz / math.ceil(abs(dx) / xStepThresh)
        stepsize = dzout / dz
        nz = math.floor(round((nz - 1) / stepsize * 100000) / 100000) + 1
        dz = dzout
        dx = math.cos(theta) * dz / xy
        zAniso = math.sin(abs(theta)) * dz / xy
    if not reverse:
        xshift, xstp = -dx, dx
    else:
        xshift, xstp = dx + math.ceil((nz - 1) * dx), -dx
    nxDs = math.ceil((nz - 1) * dx) + nx
    if objectiveScan:
        nxDs = nx
        dsS = np.eye(4)
    else:
        dsS = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [xstp, 0, 1, 0], [xshift, 0, 0, 1]], dtype=np.float64)
    if reverse:
        theta = -theta
    c = (np.array([ny, nxDs, nz], dtype=np.float64) + 1) / 2
    T1 = np.eye(4)
    T1[3, :3] = -c[[1, 0, 2]]
    S = np.diag([1.0, 1.0, zAniso, 1.0])
    ct, st = math.cos(theta), math.sin(theta)
    R = np.array([[ct, 0, -st, 0], [0, 1, 0, 0], [st, 0, ct, 0], [0, 0, 0, 1]], dtype=np.float64)
    T2 = np.eye(4)
    T2[3, :3] = (np.array(out, dtype=np.float64)[[1, 0, 2]] + 1) / 2
    T = dsS @ (T1 @ S @ R @ T2)  # imwarp's forward matrix (:210), 1-based [x y z 1] row vectors
    off = np.zeros((4, 4))
    off[3, :3] = 1
    if not objectiveScan and reverse:
        dsS[3, 0] -= dx
    M = dsS @ ((T1 + off) @ S @ R @ (T2 - off))
    tmat = np.linalg.inv(M.T)[[1, 0, 2, 3]][:, [1, 0, 2, 3]]
    if resample is not None and np.any(np.asarray(resample, dtype=np.float64) != 1):
        rs = np.asarray(resample, dtype=np.float64).reshape(3)
        RT1 = np.eye(4)
        RT1[3, :3] = -(np.array(out, dtype=np.float64)[[1, 0, 2]] + 1) / 2
        RS = np.diag([1 / rs[0], 1 / rs[1], 1 / rs[2], 1.0])
        out = [_mround(out[0] / rs[1]), _mround(out[1] / rs[0]), _mround(out[2] / rs[2])]
        RT2 = np.eye(4)
        RT2[3, :3] = (np.array(out, dtype=np.float64)[[1, 0, 2]] + 1) / 2
        T = T @ (RT1 @ RS @ RT2)
    return tmat, tuple(out), xstep0, stepsize, T


def row_structured(tmat, tol=1e-9):
    """True when the backward matrix keeps dim-1 rows (source dim 1 = output
    dim 1, source dim 2 a function of output dim 3 only, source dim 3 of
    output dims 2 and 3): the form volume_deskew_rotate_warp_mex evaluates."""
    t = np.asarray(tmat)
    return (abs(t[0][0] - 1) < tol and abs(t[0][1]) < tol and abs(t[0][2]) < tol and abs(t[0][3]) < tol
            and abs(t[1][0]) < tol and abs(t[1][1]) < tol and abs(t[2][0]) < tol)


def general_warp(vol, T, out, origin=(0, 0, 0)):
    """imwarp(vol, affine3d(T), 'linear', 'FillValues', 0, 'OutputView',
    imref3d(out, 1, 1, 1)): output voxel [x y z] (1-based, x = dim 2) reads
    the input at [x y z 1] * inv(T), trilinear over one voxel of zero padding.
    origin (x, y, z): the output window's offset (a bbox window of RA)."""
    from scipy import ndimage as ndi
    v = np.asarray(vol, dtype=np.float64)
    inv = np.linalg.inv(T)
    Y, X, Z = np.meshgrid(np.arange(1, out[0] + 1) + origin[1], np.arange(1, out[1] + 1) + origin[0],
                          np.arange(1, out[2] + 1) + origin[2], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)], axis=1) @ inv
    coords = np.stack([pts[:, 1] - 1, pts[:, 0] - 1, pts[:, 2] - 1])
    r = ndi.map_coordinates(v, coords, order=1, mode="grid-constant", cval=0.0, prefilter=False)
    return r.reshape(out).astype(np.float32)


def keys_weights(t):
    """Cubic convolution weights (Keys 1981, a = -0.5) of the samples at
    offsets -1, 0, +1, +2 from floor(x), t = x - floor(x), from the kernel
    W(s) = 1.5|s|^3 - 2.5|s|^2 + 1 (|s| <= 1), -0.5|s|^3 + 2.5|s|^2 - 4|s| + 2
    (1 < |s| < 2), 0 otherwise."""
    def W(s):
        s = abs(s)
        if s <= 1:
            return 1.5 * s ** 3 - 2.5 * s ** 2 + 1
        if s < 2:
            return -0.5 * s ** 3 + 2.5 * s ** 2 - 4 * s + 2
        return 0.0
    return [W(1 + t), W(t), W(1 - t), W(2 - t)]


def padded_extended(vol):
    """The grid imwarp's cubic resampling reads: vol padded with one zero
    voxel per side (1-based indices 0 .. n + 1), then extended by one more per
    side with Keys' boundary rule f(-1) = 3 f(0) - 3 f(1) + f(2) and
    f(n + 2) = 3 f(n + 1) - 3 f(n) + f(n - 1), axis by axis.  Returned array
    index e holds 1-based index e - 1 (e = 0 .. n + 3 for -1 .. n + 2)."""
    g = np.pad(np.asarray(vol, np.float64), 1)
    for ax in range(3):
        g = np.moveaxis(g, ax, 0)
        lo = 3 * g[0] - 3 * g[1] + g[2]
        hi = 3 * g[-1] - 3 * g[-2] + g[-3]
        g = np.concatenate([lo[None], g, hi[None]], axis=0)
        g = np.moveaxis(g, 0, ax)
    return g


def cubic_warp(vol, inv, out, origin=(0, 0, 0)):
    """imwarp(vol, tform, 'cubic', 'FillValues', 0, 'OutputView', RA) with
    inv = inv(T) (row-vector convention) and RA's window starting at 1-based
    world offset origin + 1 ([x y z] = [col row page]): tricubic convolution
    over padded_extended(vol) for points inside the padded grid [0, n + 1] on
    every axis, 0 elsewhere.  FP64 throughout."""
    v = np.asarray(vol, np.float64)
    n = np.array(v.shape)
    g = padded_extended(v)
    Y, X, Z = np.meshgrid(np.arange(1, out[0] + 1) + origin[1], np.arange(1, out[1] + 1) + origin[0],
                          np.arange(1, out[2] + 1) + origin[2], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)], axis=1) @ inv
    c = np.stack([pts[:, 1], pts[:, 0], pts[:, 2]], axis=1)  # (dim1, dim2, dim3), 1-based
    inside = np.all((c >= 0) & (c <= n + 1), axis=1)
    c = c[inside]
    fl = np.floor(c).astype(np.int64)
    t = c - fl
    w = [np.array([keys_weights(q) for q in t[:, a]]).reshape(-1, 4) for a in range(3)]  # (M, 4) per axis
    acc = np.zeros(len(c))
    for dz in range(4):
        for dx in range(4):
            for dy in range(4):
                wt = w[0][:, dy] * w[1][:, dx] * w[2][:, dz]
                live = wt != 0.0  # the taps past the extension at x = n + 1 exactly carry weight 0
                # 1-based tap fl - 1 + d sits at index fl + d of the extended grid
                i0 = np.where(live, fl[:, 0] + dy, 0)
                i1 = np.where(live, fl[:, 1] + dx, 0)
                i2 = np.where(live, fl[:, 2] + dz, 0)
                acc += np.where(live, wt * g[i0, i1, i2], 0.0)
    res = np.zeros(X.size)
    res[inside] = acc
    return res.reshape(out)


def deskew_rotate_general(vol, angle, dz, xy, reverse=False, objectiveScan=False, xStepThresh=2.0, bbox=None,
                          resample=None, interp="linear", round_interp=False):
    """deskewRotateFrame3D's imwarp form (deskewRotateFrame3D.m:73, 98-110,
    183-210), taken for 'cubic' or a resampleFactor other than ones:
    skewed-space interpolation when the step is large (round_interp: its
    result rounded as uint16, the reference's uint16 output with save16bit),
    then imwarp of the whole forward transform (with the resampling) over RA,
    the bbox window or outSize; trilinear (general_warp) or cubic
    (cubic_warp).  Linear: single-precision result; cubic: FP64."""
    tmat, out, xstep, stepsize, T = dsr_geometry(vol.shape, angle, dz, xy, reverse, objectiveScan, xStepThresh,
                                                 resample)
    v = np.asarray(vol, np.float32)
    if stepsize != 1:
        v = skewed_interp(v, xstep, stepsize, reverse)
        if round_interp:
            v = np.clip(np.floor(v.astype(np.float64) + 0.5), 0, 65535).astype(np.float32)
    if bbox is None:
        bbox = [1, 1, 1, out[0], out[1], out[2]]
    od = (bbox[3] - bbox[0] + 1, bbox[4] - bbox[1] + 1, bbox[5] - bbox[2] + 1)
    origin = (bbox[1] - 1, bbox[0] - 1, bbox[2] - 1)
    if interp == "cubic":
        return cubic_warp(v, np.linalg.inv(T), od, origin=origin)
    return general_warp(v, T, od, origin=origin)


def deskew_rotate_cubic(vol, angle, dz, xy, reverse=False, objectiveScan=False, xStepThresh=2.0, bbox=None,
                        round_interp=False, resample=None):
    """deskewRotateFrame3D(vol, ..., 'interpMethod', 'cubic'): deskew_rotate_general
    with cubic resampling.  FP64 result."""
    return deskew_rotate_general(vol, angle, dz, xy, reverse, objectiveScan, xStepThresh, bbox, resample, "cubic",
                                 round_interp)


def skewed_interp(vol, xstep, stepsize, reverse=False):
    """skewed_space_interp_defined_stepsize(im, xstep, stepsize, 'Reverse', reverse)
    (skewed_space_interp_defined_stepsize.m:24-100) in single precision: slice
    zint at zout = 1 + (zint - 1) * stepsize blends slice floor(zout) sampled at
    x + frac(zout) * xstep and the next slice at x - (1 - frac(zout)) * xstep
    (signs swapped when reversed), each by linear interpolation over zero
    padding; slices falling on an input slice are copied."""
    im = np.asarray(vol, dtype=np.float32)
    ny, nx, nz = im.shape
    if stepsize == 1:
        return im.copy()
    xa = math.ceil(xstep)
    n = math.floor(round((nz - 1) / stepsize * 100000) / 100000) + 1
    out = np.zeros((ny, nx, n), np.float32)
    for zi in range(n):
        zout = 1 + zi * stepsize
        zs = math.floor(zout)
        zw = zout - zs
        if zs >= nz:
            out[:, :, zi] = im[:, :, nz - 1]
            continue
        if reverse:
            szm = xa - zw * xstep + 1
            tzm = szm - (xa - xstep)
        else:
            szm = zw * xstep + 1
            tzm = szm + (xa - xstep)
        s, t = math.floor(szm), math.floor(tzm)
        sw, tw = szm - s, tzm - t
        if (reverse and sw == 0) or (not reverse and tw == 0):
            out[:, :, zi] = im[:, :, zs - 1]
            continue
        pad_s = np.zeros((ny, nx + xa + 1), np.float32)
        pad_t = np.zeros((ny, nx + xa + 1), np.float32)
        if reverse:
            pad_s[:, xa:xa + nx] = im[:, :, zs - 1]
            pad_t[:, 0:nx] = im[:, :, zs]
        else:
            pad_s[:, 0:nx] = im[:, :, zs - 1]
            pad_t[:, xa:xa + nx] = im[:, :, zs]
        sq = pad_s[:, s - 1:s - 1 + nx] * np.float32(1 - sw) + pad_s[:, s:s + nx] * np.float32(sw)
        tq = pad_t[:, t - 1:t - 1 + nx] * np.float32(1 - tw) + pad_t[:, t:t + nx] * np.float32(tw)
        out[:, :, zi] = sq * np.float32(1 - zw) + tq * np.float32(zw)
    return out


def row_warp(vol, tmat, bbox):
    """volume_deskew_rotate_warp_mex.cpp:217-282: output (x, y, z) over the
    1-based inclusive bbox reads source row (ty, tz), ty = t12 z + t13, tz =
    t21 y + t22 z + t23 (0-based, z / y the output indices plus the bbox start),
    blending the four rows around it bilinearly; rows with ty or tz outside
    [0, n - 1) stay zero, except ty (tz) exactly n - 1, which reads the last row."""
    v = np.asarray(vol, dtype=np.float32)
    sx, sy, sz = v.shape
    b = [int(x) for x in bbox]
    ox, oy, oz = b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1
    out = np.zeros((ox, oy, oz), np.float32)
    x0 = b[0] - 1
    for z in range(oz):
        ty = tmat[1][2] * (z + b[2] - 1) + tmat[1][3]
        if ty < 0 or ty + 1 >= sy:
            if ty + 1 != sy:
                continue
            y0, wy = sy - 2, np.float32(1.0)
        else:
            y0 = math.floor(ty)
            wy = np.float32(ty - y0)
        for y in range(oy):
            tz = tmat[2][1] * (y + b[1] - 1) + tmat[2][2] * (z + b[2] - 1) + tmat[2][3]
            if tz < 0 or tz + 1 >= sz:
                if tz + 1 != sz:
                    continue
                z0, wz = sz - 2, np.float32(1.0)
            else:
                z0 = math.floor(tz)
                wz = np.float32(tz - z0)
            one = np.float32(1)
            w00, w01 = (one - wy) * (one - wz), (one - wy) * wz
            w10, w11 = wy * (one - wz), wy * wz
            a = v[x0:x0 + ox, y0, z0]
            bb = v[x0:x0 + ox, y0, z0 + 1]
            c = v[x0:x0 + ox, y0 + 1, z0]
            d = v[x0:x0 + ox, y0 + 1, z0 + 1]
            out[:, y, z] = ((a * w00 + bb * w01) + c * w10) + d * w11
    return out


def deskew_rotate(vol, angle, dz, xy, reverse=False, objectiveScan=False, xStepThresh=2.0, bbox=None):
    """Single-precision result of deskewRotateFrame3D(single(vol), angle, dz,
    xy, 'reverse', reverse, 'objectiveScan', objectiveScan, 'xStepThresh',
    xStepThresh, 'bbox', bbox) on the linear path."""
    tmat, out, xstep, stepsize, T = dsr_geometry(vol.shape, angle, dz, xy, reverse, objectiveScan, xStepThresh)
    v = skewed_interp(vol, xstep, stepsize, reverse) if stepsize != 1 else np.asarray(vol, np.float32)
    if not row_structured(tmat):  # objective scans: the reference's imwarp form (:210)
        if bbox is not None:
            raise ValueError("bbox with an objective-scan transform is not restated here")
        return general_warp(v, T, out)
    return row_warp(v, tmat, bbox if bbox is not None else [1, 1, 1, out[0], out[1], out[2]])

"""rotateFrame3D's transform and output size on the CPU (rotateFrame3D.m:
26-100): sizes by hand for a stage-scan volume with and without the automatic
crop, the objective-scan branch, resampling; and the resampling rule the device
kernel follows, pinned against scipy's map_coordinates (linear, zero outside,
'grid-constant') at a 90-degree rotation, where it is a pure axis swap."""
import math

import numpy as np
import pytest
from scipy import ndimage

from petakit5d_amd.rotate import rotate_transform


def _warp_ref(vol, T, out):
    """imwarp 'linear' with fill 0 through map_coordinates: output voxel (r, c, p)
    samples the input at [x y z 1] = [c+1 r+1 p+1 1] inv(T) (1-based)."""
    inv = np.linalg.inv(T)
    r, c, p = np.meshgrid(np.arange(out[0]), np.arange(out[1]), np.arange(out[2]), indexing="ij")
    uvw = np.stack([c + 1.0, r + 1.0, p + 1.0, np.ones_like(r, dtype=float)], axis=-1)
    xyz = uvw @ inv
    coords = [xyz[..., 1] - 1, xyz[..., 0] - 1, xyz[..., 2] - 1]  # row, column, page (0-based)
    return ndimage.map_coordinates(np.asarray(vol, np.float64), coords, order=1, mode="grid-constant", cval=0.0)


def test_sizes_by_hand():
    ny, nx, nz, zx = 20, 100, 40, 2.0
    th = 32.45 * math.pi / 180
    full = [ny, round(nx * math.cos(th) + nz * zx / math.sin(th)), round(nz * zx * math.cos(th) + nx * math.sin(th))]
    _, od = rotate_transform((ny, nx, nz), 32.45, zx, Crop=False, ObjectiveScan=True)
    assert list(od) == full
    # objective scan: never cropped
    _, od = rotate_transform((ny, nx, nz), 32.45, zx, Crop=True, ObjectiveScan=True)
    assert list(od) == full
    # resample by 2 in every axis halves each size (rounded)
    _, od2 = rotate_transform((ny, nx, nz), 32.45, zx, Crop=False, ObjectiveScan=True, resample=[2, 2, 2])
    assert list(od2) == [int(math.floor(v / 2 + 0.5)) for v in full]


def test_stage_scan_crop_from_projection():
    """A deskewed stage-scan volume occupies a parallelogram in x-z; the crop
    takes its extent across / along the rotated axis minus 4 (rotateFrame3D.m:46-66)."""
    ny, nx, nz, zx = 4, 120, 30, 1.0
    proj = np.zeros((nx, nz))
    for k in range(nz):  # each plane's data shifted by 2 columns per plane, 40 wide
        proj[2 * k:2 * k + 40, k] = 1
    th = 32.45 * math.pi / 180
    a, b = math.sin(-th), math.cos(-th)
    d = np.diff((proj != 0).T.astype(int), axis=1)
    z, x = np.nonzero(d)[0] + 1.0, np.nonzero(d)[1] + 1.0
    c = a * x + b * z * zx
    h = math.floor(c.max() - c.min() + 0.5) - 4
    cx = b * x - a * z * zx
    w = math.floor(cx.max() - cx.min() + 0.5) - 4
    assert h < max(abs(a * (nx - 1) + b * (nz - 1) * zx), abs(a * (nx - 1) + b * (1 - nz) * zx))
    _, od = rotate_transform((ny, nx, nz), 32.45, zx, Crop=True, proj=proj)
    assert od == (ny, w, h)
    _, od = rotate_transform((ny, nx, nz), 32.45, zx, Crop=False, proj=proj)
    assert od[1] == round(nx * math.cos(th) + nz * zx / math.sin(th))


def test_ninety_degrees_swaps_x_and_z():
    """At 90 degrees (zxRatio 1) output column u holds input plane z = u and
    output plane w holds input column x = nx + 1 - w."""
    g = np.random.default_rng(0)
    vol = g.random((5, 7, 9))
    T, od = rotate_transform(vol.shape, 90.0, 1.0, Crop=False, ObjectiveScan=True)
    assert od == (5, 9, 7)
    out = _warp_ref(vol, T, od)
    exp = np.transpose(vol, (0, 2, 1))[:, :, ::-1]
    np.testing.assert_allclose(out, exp, rtol=0, atol=1e-9)


@pytest.mark.parametrize("angle", [32.45, -20.0, 57.3])
def test_transform_closed_forms_general_angle(angle):
    """At a general angle: T maps the input centre ((n+1)/2 per axis) onto the
    output centre; every corner of the input box lands inside the uncropped
    output box; reverse=True is the x-mirror of the forward map on both sides
    (R(-theta) = M R(theta) M with M the x mirror, which commutes with the z
    scaling): T_rev(p) = mirror_out(T_fwd(mirror_in(p)))."""
    shape, zx = (16, 90, 40), 1.7
    ny, nx, nz = shape
    T, od = rotate_transform(shape, angle, zx, Crop=False, ObjectiveScan=True)
    cin = np.array([(nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2, 1.0])
    cout = np.array([(od[1] + 1) / 2, (od[0] + 1) / 2, (od[2] + 1) / 2])
    np.testing.assert_allclose((cin @ T)[:3], cout, rtol=0, atol=1e-9)
    corners = np.array([[x, y, z, 1.0] for x in (0.5, nx + 0.5) for y in (0.5, ny + 0.5) for z in (0.5, nz + 0.5)])
    q = corners @ T
    lim = np.array([od[1], od[0], od[2]]) + 0.5
    assert np.all(q[:, :3] >= 0.5 - 1.0) and np.all(q[:, :3] <= lim + 1.0)  # within rounding of the box
    Tr, odr = rotate_transform(shape, angle, zx, reverse=True, Crop=False, ObjectiveScan=True, outSize=od)
    assert odr == od
    g = np.random.default_rng(1)
    p = np.c_[g.uniform(1, nx, 50), g.uniform(1, ny, 50), g.uniform(1, nz, 50), np.ones(50)]
    pm = p.copy()
    pm[:, 0] = nx + 1 - p[:, 0]
    fwd = pm @ T
    fwd[:, 0] = od[1] + 1 - fwd[:, 0]
    np.testing.assert_allclose(p @ Tr, fwd, rtol=0, atol=1e-9)

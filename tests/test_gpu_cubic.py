"""'cubic' resampling on the device (pk_affine_warp3d_cubic,
pk_deskew_rotate_frame3d_cubic) against the oracle's FP64 restatement of
imwarp's cubic convolution (oracle/deskew_oracle.py: cubic_warp,
deskew_rotate_cubic): rotateFrame3D, deskewFrame3D and deskewRotateFrame3D
with 'cubic', forward and reversed, with and without skewed-space
interpolation, uint16 with save16bit, a bbox, an objective scan; integer
shifts reproduce samples bit for bit."""
import math

import numpy as np
import pytest

import deskew_oracle as D

pytestmark = pytest.mark.gpu


def _smooth(shape, seed):
    from scipy import ndimage as ndi
    g = np.random.default_rng(seed)
    v = ndi.gaussian_filter(g.random(shape) * 4000, 1.5) + g.random(shape) * 50
    return np.asfortranarray(v.astype(np.float32))


def _relmax(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("reverse", [False, True])
def test_rotate_cubic_vs_oracle(reverse):
    import petakit5d_amd as pk
    from petakit5d_amd.rotate import rotate_transform
    print(f"[cubic] rotateFrame3D reverse={reverse}", flush=True)
    v = _smooth((24, 60, 30), 1)
    za = math.sin(math.radians(32.45)) * 0.3 / 0.108
    got = pk.rotateFrame3D(v, 32.45, za, reverse, Interp="cubic")
    T, od = rotate_transform(v.shape, 32.45, za, reverse, proj=v.max(axis=0))
    ref = D.cubic_warp(v, np.linalg.inv(T), od)
    assert got.shape == tuple(od)
    e = _relmax(got, ref)
    print(f"  out {got.shape} max rel err {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("reverse,crop", [(False, False), (True, False), (False, True)])
def test_deskew_cubic_vs_oracle(reverse, crop):
    import petakit5d_amd as pk
    from petakit5d_amd.deskew import deskew_transform
    v = _smooth((20, 48, 36), 2)
    got = pk.deskewFrame3D(v, 32.45, 0.3, 0.108, reverse, Crop=crop, Interp="cubic")
    S, od = deskew_transform(v.shape, 32.45, 0.3, 0.108, reverse, crop)
    ref = D.cubic_warp(v, np.linalg.inv(S), od)
    assert got.shape == tuple(od)
    assert _relmax(got, ref) < 1e-5


def test_integer_shift_cubic_equals_linear_bitwise():
    """A shear by exactly 2 columns per plane (angle 0, dz / xy = 2): both
    interpolations read samples (or the zero pad) only, so cubic and linear
    give the same bits."""
    import petakit5d_amd as pk
    v = _smooth((16, 30, 12), 3)
    lin = pk.deskewFrame3D(v, 0.0, 0.5, 0.25, False)
    cub = pk.deskewFrame3D(v, 0.0, 0.5, 0.25, False, Interp="cubic")
    np.testing.assert_array_equal(cub, lin)


@pytest.mark.parametrize("dz,reverse", [(0.5, False), (0.5, True), (0.2, False), (0.2, True)])
def test_deskew_rotate_cubic_vs_oracle(dz, reverse):
    """dz 0.5: dx = 3.91 > 2, skewed-space interpolation first; dz 0.2: none."""
    import petakit5d_amd as pk
    print(f"[cubic] deskewRotateFrame3D dz={dz} reverse={reverse}", flush=True)
    v = _smooth((12, 72, 30), 4)
    got = pk.deskewRotateFrame3D(v, 32.45, dz, 0.108, reverse=reverse, interpMethod="cubic")
    ref = D.deskew_rotate_cubic(v, 32.45, dz, 0.108, reverse=reverse)
    assert got.shape == ref.shape and got.dtype == np.float32
    e = _relmax(got, ref)
    print(f"  out {got.shape} max rel err {e:.2e}", flush=True)
    assert e < 1e-5


def test_deskew_rotate_cubic_uint16_bbox_objective():
    import petakit5d_amd as pk
    v16 = np.asfortranarray(_smooth((10, 64, 24), 5).astype(np.uint16))
    # uint16 with save16bit and skewed-space interpolation: the interpolated slices are
    # uint16 values (the reference's interpolation output), the result uint16
    got = pk.deskewRotateFrame3D(v16, 32.45, 0.5, 0.108, save16bit=True, interpMethod="cubic")
    ref = D.deskew_rotate_cubic(v16, 32.45, 0.5, 0.108, round_interp=True)
    assert got.dtype == np.uint16
    d = np.abs(got.astype(np.int64) - np.clip(np.floor(ref + 0.5), 0, 65535).astype(np.int64))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    # uint16 without interpolation: imwarp keeps the class, uint16 even without save16bit
    got2 = pk.deskewRotateFrame3D(v16, 32.45, 0.2, 0.108, interpMethod="cubic")
    assert got2.dtype == np.uint16
    # a bbox is the output window of the same warp
    vf = _smooth((10, 64, 24), 6)
    full = D.deskew_rotate_cubic(vf, 32.45, 0.3, 0.108)
    box = [2, 5, 3, 9, full.shape[1] - 6, full.shape[2] - 2]
    gb = pk.deskewRotateFrame3D(vf, 32.45, 0.3, 0.108, bbox=box, interpMethod="cubic")
    rb = full[box[0] - 1:box[3], box[1] - 1:box[4], box[2] - 1:box[5]]
    assert gb.shape == rb.shape and _relmax(gb, rb) < 1e-5
    # objective scan
    vo = _smooth((12, 40, 20), 7)
    go = pk.deskewRotateFrame3D(vo, 20.0, 0.3, 0.1, objectiveScan=True, interpMethod="cubic")
    ro = D.deskew_rotate_cubic(vo, 20.0, 0.3, 0.1, objectiveScan=True)
    assert go.shape == ro.shape and _relmax(go, ro) < 1e-5

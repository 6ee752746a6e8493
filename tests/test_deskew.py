"""deskewFrame3D geometry and the oracle's resampling, on the CPU: output
sizes from the C-ABI's host-only pk_deskew_out_dims equal the oracle's for
forward / reverse / cropped geometries, and the oracle is pinned by closed
forms -- integer shifts move planes exactly, linear interpolation reproduces a
ramp along x, the padded edge blends the last column with zero."""
import math

import numpy as np
import pytest

import deskew_oracle as D
from petakit5d_amd.deskew import deskew_out_size


@pytest.mark.parametrize("angle,dz,xy", [(32.45, 0.3, 0.108), (31.5, 0.5, 0.1), (0.0, 0.2, 0.1), (60.0, 0.4, 0.1)])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("crop", [False, True])
@pytest.mark.parametrize("shape", [(40, 64, 30), (8, 10, 120), (5, 3, 1)])
def test_out_size_matches_oracle(angle, dz, xy, reverse, crop, shape):
    nx_out, _, _ = D.geometry(shape, angle, dz, xy, reverse, crop)
    assert deskew_out_size(shape, angle, dz, xy, reverse, crop) == (shape[0], nx_out, shape[2])


def test_out_size_by_hand():
    # dx = 0.3 cos(32.45 deg) / 0.108 = 2.3445...; ceil(29 dx) = 68
    assert deskew_out_size((40, 64, 30), 32.45, 0.3, 0.108) == (40, 68 + 64, 30)
    # crop: 132 >= 128 -> floor(64 / 2) = 32 per side
    assert deskew_out_size((40, 64, 30), 32.45, 0.3, 0.108, Crop=True) == (40, 132 - 64, 30)
    # crop below 2 nx: floor(68 / 2) = 34 per side
    assert deskew_out_size((40, 100, 30), 32.45, 0.3, 0.108, Crop=True) == (40, 168 - 68, 30)


@pytest.mark.parametrize("reverse", [False, True])
def test_integer_shift_moves_planes(reverse):
    """dx = 0.2 / 0.1 = 2 exactly at angle 0: plane p moves 2 (p - 1) columns
    (2 (nz - p) reversed) with no interpolation."""
    g = np.random.default_rng(1)
    vol = g.random((6, 9, 5)).astype(np.float32)
    out = D.deskew(vol, 0.0, 0.2, 0.1, reverse=reverse)
    assert out.shape == (6, 9 + 8, 5)
    exp = np.zeros_like(out)
    for p in range(5):
        s = 2 * (4 - p) if reverse else 2 * p
        exp[:, s:s + 9, p] = vol[:, :, p]
    np.testing.assert_array_equal(out, exp)


def test_linear_ramp_and_edge_blend():
    """A ramp vol(:, x, :) = x is reproduced at every interior position x of
    every plane; just past the last column the value falls linearly to 0."""
    ny, nx, nz = 3, 20, 7
    vol = np.broadcast_to(np.arange(1, nx + 1, dtype=np.float32)[None, :, None], (ny, nx, nz))
    angle, dz, xy = 31.5, 0.5, 0.1
    out = D.deskew(vol, angle, dz, xy)
    nx_out, xstep, xshift = D.geometry((ny, nx, nz), angle, dz, xy)
    for p in range(1, nz + 1):
        x = np.arange(1, nx_out + 1) - xstep * p - xshift
        inside = (x >= 1) & (x <= nx)
        np.testing.assert_allclose(out[0, inside, p - 1], x[inside], rtol=0, atol=1e-9)
        blend = (x > nx) & (x < nx + 1)
        np.testing.assert_allclose(out[0, blend, p - 1], nx * (nx + 1 - x[blend]), rtol=0, atol=1e-9)
        lead = (x > 0) & (x < 1)
        np.testing.assert_allclose(out[0, lead, p - 1], x[lead], rtol=0, atol=1e-9)  # 0 blended with vol(1) = 1
        assert np.all(out[0, (x < 0) | (x > nx + 1), p - 1] == 0)


def test_geometry_constants():
    dx = math.cos(32.45 * math.pi / 180) * 0.3 / 0.108
    nx_out, xstep, xshift = D.geometry((4, 10, 11), 32.45, 0.3, 0.108, reverse=True)
    assert (xstep, xshift) == (-dx, dx + math.ceil(10 * dx)) and nx_out == math.ceil(10 * dx) + 10


# ---------------------------------------------------------------------------------------------------------------------
# deskewRotateFrame3D (deskewRotateFrame3D.m:50-253): the oracle's geometry and resampling against closed forms.
@pytest.mark.parametrize("angle,dz,xy", [(32.45, 0.3, 0.108), (32.45, 0.2, 0.108), (31.5, 0.5, 0.1)])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("shape", [(40, 64, 30), (16, 100, 60)])
def test_dsr_out_size_matches_oracle(angle, dz, xy, reverse, shape):
    from petakit5d_amd.deskew_rotate import deskew_rotate_out_size
    _, out, _, _, _ = D.dsr_geometry(shape, angle, dz, xy, reverse)
    assert deskew_rotate_out_size(shape, angle, dz, xy, reverse) == out


def test_dsr_geometry_by_hand():
    """dx = 0.3 cos(32.45 deg) / 0.108 = 2.344 > 2: |dx| / 2 < 1.5, so dz' =
    dz / ceil(dz / (0.108 * 2 / cos)) = 0.15 (step 0.5, 2 * 255 + 1 slices);
    outSize = round([ny, 511 cos + 255 zAniso / sin, 511 sin - 4]) with
    zAniso = sin * 0.3 / 0.108."""
    t = math.radians(32.45)
    tm, out, xs, step, _ = D.dsr_geometry((512, 512, 256), 32.45, 0.3, 0.108)
    za = math.sin(t) * 0.3 / 0.108
    assert out == (512, round(511 * math.cos(t) + 255 * za / math.sin(t)), round(511 * math.sin(t) - 4))
    assert abs(xs - math.cos(t) * 0.3 / 0.108) < 1e-12 and step == 0.5
    assert D.row_structured(tm)
    _, _, _, step2, _ = D.dsr_geometry((512, 512, 256), 32.45, 0.2, 0.108)
    assert step2 == 1.0  # dx = 1.56 <= 2: no interpolation


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("xstep,step", [(2.344, 0.5), (3.1, 1 / 3), (2.7, 0.25)])
def test_skewed_interp_follows_the_skew(reverse, xstep, step):
    """Slices of a stage scan are shifted copies: slice z (1-based) holds
    G(x + s (z - 1) xstep), s = +1 (-1 reversed), the deskew convention of
    deskewFrame3D.m (x' = x + xstep (z - 1)).  With G linear, every
    interpolated slice at zout must equal G(x + s (zout - 1) xstep) wherever
    its samples stay inside the slices (linear interpolation is exact there)."""
    ny, nx, nz = 3, 60, 7
    s = -1 if reverse else 1
    G = lambda u: 0.75 * u + 2.0  # noqa: E731
    x = np.arange(nx, dtype=np.float64)
    vol = np.stack([np.broadcast_to(G(x + s * z * xstep), (ny, nx)) for z in range(nz)], axis=2).astype(np.float32)
    out = D.skewed_interp(vol, xstep, step, reverse)
    assert out.shape[2] == round((nz - 1) / step) + 1
    m = int(np.ceil(xstep)) + 1
    for zi in range(out.shape[2]):
        zout = 1 + zi * step
        want = G(x + s * (zout - 1) * xstep)
        np.testing.assert_allclose(out[:, m:nx - m, zi], np.broadcast_to(want[m:nx - m], (ny, nx - 2 * m)),
                                   rtol=2e-6, atol=2e-5)


def test_dsr_objective_scan_identity_and_stretch():
    """Objective scan at angle 0: no shear, no rotation.  dz = xyPixelSize is the
    identity; dz = 2 xyPixelSize doubles z, output slice Z (1-based) sampling
    input slice (Z - (2 nz + 1) / 2) / 2 + (nz + 1) / 2."""
    g = np.random.default_rng(2)
    v = g.random((6, 7, 5)).astype(np.float32)
    np.testing.assert_array_equal(D.deskew_rotate(v, 0.0, 0.1, 0.1, objectiveScan=True), v)
    r = D.deskew_rotate(v, 0.0, 0.2, 0.1, objectiveScan=True)
    assert r.shape == (6, 7, 10)
    for Z in range(1, 11):
        zs = (Z - 5.5) / 2 + 3  # 1-based source slice
        z0 = math.floor(zs)
        f = zs - z0
        lo = v[:, :, z0 - 1] if 1 <= z0 <= 5 else 0 * v[:, :, 0]
        hi = v[:, :, z0] if 1 <= z0 + 1 <= 5 else 0 * v[:, :, 0]
        np.testing.assert_allclose(r[:, :, Z - 1], (1 - f) * lo + f * hi, rtol=1e-6, atol=1e-6)


def test_dsr_equals_deskew_then_rotate_on_smooth_data():
    """deskewRotateFrame3D is deskewFrame3D followed by rotateFrame3D in one
    resampling (deskewRotateFrame3D.m:1-6): on a smooth volume the combined
    result matches the two-step one (independently pinned oracles, same output
    box) up to the extra interpolation, and a wrong sign, centre or scale
    would not.  Forward (not reversed) scans: the reference's default fast path
    builds its 0-based matrix from the 1-based shear xshift = -dx
    (deskewRotateFrame3D.m:131-133, 212-218 corrects only the reversed case),
    so its output samples the deskewed volume dx columns further along x than
    the imwarp form of :210 -- the oracle keeps that, and the two-step result
    is taken of the deskewed volume moved by dx."""
    from scipy import ndimage as ndi
    from petakit5d_amd.rotate import rotate_transform
    ny, nx, nz = 8, 96, 48
    g = np.random.default_rng(3)
    obj = np.zeros((ny, nx, nz))
    obj[4, g.integers(20, 76, 12), g.integers(10, 38, 12)] = 1000.0
    vol = ndi.gaussian_filter(obj, (1.0, 3.0, 3.0)).astype(np.float32)
    angle, dz, xy = 32.45, 0.2, 0.108  # dx = 1.56: no skewed-space interpolation
    dx = math.cos(math.radians(angle)) * dz / xy
    za = math.sin(math.radians(angle)) * dz / xy
    for reverse in (False, True):
        dsr = D.deskew_rotate(vol, angle, dz, xy, reverse=reverse)
        ds = D.deskew(vol, angle, dz, xy, reverse=reverse).astype(np.float32)
        T, od = rotate_transform(ds.shape, angle, za, reverse, Crop=False, ObjectiveScan=True, outSize=dsr.shape)
        shifted = ds if reverse else ndi.shift(ds, (0, -dx, 0), order=1, mode="constant")
        two = D.general_warp(shifted, T, od)
        err = np.linalg.norm(dsr - two) / np.linalg.norm(two)
        assert err < 0.05, (reverse, err)
        if not reverse:  # the unshifted two-step result is far off: the offset is real
            assert np.linalg.norm(dsr - D.general_warp(ds, T, od)) / np.linalg.norm(two) > 0.2


# ---------------------------------------------------------------- 'cubic'
# imwarp's 'cubic' restated as cubic convolution (Keys, a = -0.5) over the
# volume padded with one zero voxel, Keys' boundary rule past the padded grid,
# fill value 0 outside it.  Pinned by closed forms: the kernel's textbook
# weights, exact reproduction of quadratics where every tap is a sample,
# integer shifts that copy samples bit for bit, and the hand-derived edge
# values of the boundary rule.

def test_keys_weights_closed_forms():
    assert D.keys_weights(0.0) == [0.0, 1.0, 0.0, 0.0]
    assert D.keys_weights(0.5) == [-1 / 16, 9 / 16, 9 / 16, -1 / 16]  # the half-sample filter (-1, 9, 9, -1) / 16
    for t in np.linspace(0, 0.999, 37):
        w = D.keys_weights(float(t))
        assert abs(sum(w) - 1) < 1e-14
        for a, b, c in ((1.0, 2.0, 3.0), (-0.5, 0.3, 1.7)):
            f = [a + b * x + c * x * x for x in (-1, 0, 1, 2)]
            assert abs(sum(wk * fk for wk, fk in zip(w, f)) - (a + b * t + c * t * t)) < 1e-12


def test_cubic_warp_integer_shift_and_quadratic():
    v = np.random.default_rng(0).random((5, 6, 7))
    inv = np.eye(4)
    inv[3, :3] = [2, -1, 3]  # reads x + 2, y - 1, z + 3
    r = D.cubic_warp(v, inv, (5, 6, 7))
    exp = np.zeros_like(v)
    exp[1:, :4, :4] = v[:4, 2:, 3:]
    assert np.array_equal(r, exp)
    # a quadratic, rotated about y: exact wherever all 64 taps are samples
    Y, X, Z = np.meshgrid(np.arange(1, 21), np.arange(1, 23), np.arange(1, 19), indexing="ij")
    quad = lambda x, y, z: 0.3 + 0.02 * x - 0.05 * y + 0.01 * z + 0.003 * x * x - 0.002 * x * z + 0.004 * y * z \
        + 0.001 * y * y  # noqa: E731
    th = 0.3
    R = np.array([[math.cos(th), 0, -math.sin(th), 0], [0, 1, 0, 0], [math.sin(th), 0, math.cos(th), 0],
                  [0, 0, 0, 1]])
    T1, T2 = np.eye(4), np.eye(4)
    T1[3, :3] = -np.array([11.5, 10.5, 9.5])
    T2[3, :3] = [11.5, 10.5, 9.5]
    inv = np.linalg.inv(T1 @ R @ T2)
    r = D.cubic_warp(quad(X, Y, Z), inv, (20, 22, 18)).ravel()
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)], 1) @ inv
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    inner = (x >= 2) & (x < 21) & (y >= 2) & (y < 19) & (z >= 2) & (z < 17)
    assert inner.sum() > 4000
    assert np.abs(r[inner] - quad(x, y, z)[inner]).max() < 1e-12


def test_cubic_warp_edge_closed_forms():
    """Along x at integer y, z: half a sample outside the data (x = 0.5) the
    taps are f(-1) = -3 v1 + v2 (boundary rule), f(0) = 0 (pad), v1, v2, so the
    value is 0.75 v1 - 0.125 v2; at x = n + 0.5 it is 0.75 vn - 0.125 v(n-1);
    on the pad (x = 0, n + 1) it is 0, and beyond it the fill value 0."""
    v = np.random.default_rng(1).random((4, 7, 3)) + 1
    for xq, want in ((0.5, 0.75 * v[:, 0] - 0.125 * v[:, 1]), (7.5, 0.75 * v[:, 6] - 0.125 * v[:, 5]),
                     (0.0, 0 * v[:, 0]), (8.0, 0 * v[:, 0]), (-0.25, 0 * v[:, 0]), (8.3, 0 * v[:, 0])):
        inv = np.eye(4)
        inv[0, 0] = 0.0
        inv[3, 0] = xq  # every output column reads x = xq
        r = D.cubic_warp(v, inv, (4, 1, 3))
        np.testing.assert_allclose(r[:, 0, :], want, rtol=0, atol=1e-14)


def test_dsr_cubic_matches_linear_on_smooth_data():
    """The cubic and the linear deskew-rotate (independent resampling paths:
    the general imwarp form vs the row warp) agree on a smooth volume up to
    interpolation error, for a forward scan with skewed-space interpolation
    and a reversed one without; the cubic form carries the reference's dx
    offset only on its fast path, so forward scans are compared shifted."""
    from scipy import ndimage as ndi
    ny, nx, nz = 6, 80, 40
    g = np.random.default_rng(5)
    obj = np.zeros((ny, nx, nz))
    obj[3, g.integers(20, 60, 10), g.integers(8, 32, 10)] = 1000.0
    vol = ndi.gaussian_filter(obj, (1.0, 3.0, 3.0)).astype(np.float32)
    for dz, reverse in ((0.5, False), (0.2, True)):
        lin = D.deskew_rotate(vol, 32.45, dz, 0.108, reverse=reverse)
        cub = D.deskew_rotate_cubic(vol, 32.45, dz, 0.108, reverse=reverse)
        assert cub.shape == lin.shape
        if not reverse:  # the fast path's forward-scan offset: dx of the (interpolated) slices along x
            _, _, xstep, stepsize, _ = D.dsr_geometry(vol.shape, 32.45, dz, 0.108, reverse)
            ref = D.deskew_rotate_cubic(ndi.shift(vol.astype(np.float64), (0, -xstep * stepsize, 0), order=3,
                                                  mode="constant"), 32.45, dz, 0.108, reverse=reverse)
            cub = ref
        err = np.linalg.norm(cub - lin) / np.linalg.norm(lin)
        assert err < 0.08, (dz, reverse, err)


# ---------------------------------------------------------------- resampleFactor
# deskewRotateFrame3D.m:183-196: T * (RT1 * RS * RT2) into round(outSize ./ rs([2 1 3])).

@pytest.mark.parametrize("rs", [(1.0, 1.0, 1.0), (1.5, 0.5, 2.0), (2.0, 2.0, 2.0), (0.8, 1.0, 1.3)])
@pytest.mark.parametrize("reverse", [False, True])
def test_dsr_resample_geometry_closed_forms(rs, reverse):
    """Ones leave the geometry alone; otherwise the output size is
    round(outSize ./ rs([2 1 3])) (the C-ABI's host-only size agrees), the
    new box centre maps to the input point the old centre mapped to, and a
    unit step along output axis k covers rs(k) steps of the unresampled
    output (the inverse map's rows scale by rs)."""
    from petakit5d_amd.deskew_rotate import deskew_rotate_out_size
    shape = (40, 64, 30)
    _, out0, _, _, T0 = D.dsr_geometry(shape, 32.45, 0.3, 0.108, reverse)
    _, out1, _, _, T1 = D.dsr_geometry(shape, 32.45, 0.3, 0.108, reverse, resample=rs)
    assert deskew_rotate_out_size(shape, 32.45, 0.3, 0.108, reverse, resampleFactor=rs) == out1
    if rs == (1.0, 1.0, 1.0):
        assert out1 == out0 and np.array_equal(T1, T0)
        return
    assert out1 == (D._mround(out0[0] / rs[1]), D._mround(out0[1] / rs[0]), D._mround(out0[2] / rs[2]))
    c0 = np.array([(out0[1] + 1) / 2, (out0[0] + 1) / 2, (out0[2] + 1) / 2, 1.0])
    c1 = np.array([(out1[1] + 1) / 2, (out1[0] + 1) / 2, (out1[2] + 1) / 2, 1.0])
    i0, i1 = np.linalg.inv(T0), np.linalg.inv(T1)
    np.testing.assert_allclose(c1 @ i1, c0 @ i0, atol=1e-9)
    for k in range(3):
        e = np.zeros(4)
        e[k] = 1.0
        np.testing.assert_allclose(e @ i1, rs[k] * (e @ i0), atol=1e-12)


def test_dsr_resample_y_by_two_is_the_row_pair_mean():
    """Rows pass through the rotation about y untouched, so with rs = [1 2 1]
    (ny even) output row r sits at y = 2 r - 1/2 of the unresampled output:
    trilinear interpolation gives the mean of unresampled rows 2r - 1 and 2r
    exactly (both share the x / z weights)."""
    g = np.random.default_rng(11)
    vol = g.random((12, 40, 20)).astype(np.float32) * 1000
    for interp_dz in (0.2, 0.5):  # without and with skewed-space interpolation
        full = D.deskew_rotate_general(vol, 32.45, interp_dz, 0.108, resample=None).astype(np.float64)
        half = D.deskew_rotate_general(vol, 32.45, interp_dz, 0.108, resample=(1.0, 2.0, 1.0)).astype(np.float64)
        assert half.shape == (6,) + full.shape[1:]
        np.testing.assert_allclose(half, 0.5 * (full[0::2] + full[1::2]), rtol=1e-5, atol=1e-3)

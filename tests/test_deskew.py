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

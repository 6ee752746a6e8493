"""psf_gen_new (decon/psf_gen_new.m:1-140) restated with numpy / scipy.
Pinned by closed forms: a PSF with empty border slices skips the masking, so
the output is exactly the circular shift that puts MATLAB's first maximum at
round((n + 1) / 2); the axial resampling keeps the DC term (the total sum) and
halves the length for dz_data = 2 dz_psf; the masked method removes a flat
background and keeps the peak's component.  The toolbox steps' own tie and
border rules (medfilt3, bwlabeln, imclose) stay unpinned without MATLAB."""
import numpy as np

from petakit5d_amd.psfgen import psf_gen_new


def test_zero_border_psf_is_only_centred():
    p = np.zeros((21, 19, 30), np.float32)
    p[4, 6, 9] = 5.0
    p[5, 6, 9] = 5.0  # a tie: the first maximum in column-major order (y = 4) wins
    p[7, 8, 11] = 2.0
    out = psf_gen_new(p, 0.1, 0.1)
    # shift = round(([21 19 30] + 1) / 2 - [5 7 10]) = round([6 3 5.5]) = [6 3 6]
    np.testing.assert_array_equal(out, np.roll(p, (6, 3, 6), axis=(0, 1, 2)))
    assert out[10, 9, 15] == 5.0  # the peak at MATLAB's centre (11, 10, 16), 0-based (10, 9, 15)


def test_axial_resampling_keeps_the_sum_and_halves_the_length():
    g = np.arange(24, dtype=np.float64)
    zc = np.arange(40, dtype=np.float64)
    blob = (np.exp(-((g[:, None, None] - 12) ** 2 + (g[None, :, None] - 12) ** 2) / 8.0)
            * np.exp(-((zc[None, None, :] - 20) ** 2) / 18.0)).astype(np.float32)
    blob[:, :, :6] = 0
    blob[:, :, -6:] = 0  # empty border slices: no masking
    out = psf_gen_new(blob, 0.1, 0.2)
    assert out.shape == (24, 24, 20)
    assert abs(out.sum(dtype=np.float64) - blob.sum(dtype=np.float64)) < 1e-3 * blob.sum(dtype=np.float64)
    assert out.min() >= 0


def test_masked_method_removes_background_keeps_the_psf():
    rng = np.random.default_rng(3)
    g = np.arange(32, dtype=np.float64)
    zc = np.arange(41, dtype=np.float64)
    blob = 2000 * np.exp(-((g[:, None, None] - 15) ** 2 + (g[None, :, None] - 17) ** 2) / 4.0
                         - ((zc[None, None, :] - 19) ** 2) / 12.0)
    raw = (blob + 100 + rng.normal(0, 2, size=blob.shape)).astype(np.float32)
    out = psf_gen_new(raw, 0.1, 0.1, 1.5, "masked")
    assert out.min() >= 0
    pk = np.unravel_index(int(np.argmax(out.ravel(order="F"))), out.shape, order="F")
    # the peak (1-based 16, 18, 20) moves by round(([32 32 41] + 1) / 2 - peak) = round([0.5 -1.5 1])
    # = [1 -2 1], halves away from zero: to (17, 16, 21), 0-based (16, 15, 20)
    assert tuple(int(v) for v in pk) == (16, 15, 20)
    # far from the PSF the background is gone
    assert np.all(out[:4, :4, :] == 0) and np.all(out[:, :, :3] == 0)
    # the PSF's energy survives (background ~100 per voxel removed)
    core = out[12:21, 12:21, 14:27].sum(dtype=np.float64)
    assert 0.8 * blob.sum() < core < 1.1 * blob.sum()


def test_median_method_and_2d_psf_run():
    rng = np.random.default_rng(4)
    g = np.arange(30, dtype=np.float64)
    blob = 1500 * np.exp(-((g[:, None, None] - 14) ** 2 + (g[None, :, None] - 15) ** 2) / 5.0
                         - ((np.arange(25.0)[None, None, :] - 12) ** 2) / 10.0)
    raw = (blob + 50 + rng.normal(0, 1, size=blob.shape)).astype(np.float32)
    out = psf_gen_new(raw, 0.1, 0.1, 1.5, "median")
    assert out.shape == raw.shape and out.min() >= 0 and out.max() > 0
    out2 = psf_gen_new(raw[:, :, 12], 0.1, 0.1, 1.5, "masked")
    assert out2.shape == (30, 30, 1) and out2.min() >= 0 and out2.max() > 0

"""psf_gen_new (decon/psf_gen_new.m:1-140) restated with numpy / scipy.
Pinned by closed forms: a PSF with empty border slices skips the masking, so
the output is exactly the circular shift that puts MATLAB's first maximum at
round((n + 1) / 2); the axial resampling keeps the DC term (the total sum) and
halves the length for dz_data = 2 dz_psf; the masked method removes a flat
background and keeps the peak's component.  The toolbox steps are pinned by
hand-derived cases of their documented rules: medfilt3's symmetric and
medfilt2's zero padding, 26-connectivity, imclose's false / true border
padding, and strel('disk', 3) as the decomposed 5 x 5 element."""
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


# ------------------------------------------------------------------ toolbox steps, pinned by hand
def test_median_filter_border_rules():
    """medfilt3 pads symmetrically, medfilt2 with zeros (their documented
    defaults).  Corner of a 2x2x2 volume v = 4i + 2j + k: the symmetric
    3x3x3 window holds v000 x8, v001 / v010 / v100 x4, v011 / v101 / v110 x2,
    v111 x1, whose 14th value is 2 (zero padding would give 0).  Corner of the
    2x2 image [[1 2] [3 4]]: the zero-padded window is five zeros and 1..4,
    median 0 (symmetric padding would give 2)."""
    from petakit5d_amd.psfgen import _medfilt2, _medfilt3
    v = np.arange(8, dtype=np.float64).reshape(2, 2, 2)
    assert _medfilt3(v)[0, 0, 0] == 2.0
    assert _medfilt2(np.array([[1.0, 2.0], [3.0, 4.0]]))[0, 0] == 0.0


def test_strel_disk_of_the_2d_branch():
    """imclose(BW, strel('disk', 3)) in the 2-D branch uses the decomposed
    disk (default N = 4): for r = 3 the four periodic lines sum to a 3x3
    square and the compensating lines make it 5x5 -- not the 7x7 exact disk.
    The product's construction equals the oracle's (pinned against the
    documented r = 5 element and brute-force Minkowski sums) for r <= 25."""
    import pk_oracle
    from petakit5d_amd.psfgen import _strel_disk
    assert np.array_equal(_strel_disk(3), np.ones((5, 5), bool))
    for r in range(26):
        np.testing.assert_array_equal(_strel_disk(r), pk_oracle.strel_disk(r))


def test_component_mask_connectivity_and_closing_borders():
    """bwareaopen / bwlabeln with 26-connectivity: two voxels touching only at
    a corner form one component (kept at P = 2; 6-connectivity would drop
    both).  imclose pads the dilation with false and the erosion with true, so
    a mask filling the image stays whole (a zero-padded erosion would eat two
    pixels off every border) and, in a U-shaped component, a notch 3 pixels
    wide fills under the 5x5 element while one 5 pixels wide does not."""
    from petakit5d_amd.psfgen import _component_mask, _sphere, _strel_disk
    m = np.zeros((6, 6, 6))
    m[2, 2, 2] = 5.0
    m[3, 3, 3] = 1.0
    bw = _component_mask(m, m, 2, True, np.ones((1, 1, 1), bool))
    assert bw[2, 2, 2] and bw[3, 3, 3] and bw.sum() == 2
    full = np.ones((9, 9))
    assert _component_mask(full, full, 1, True, _strel_disk(3)).all()
    for gap, closes in ((3, True), (5, False)):
        img = np.zeros((13, 20 + gap))
        img[3:8, 2:10] = 1.0
        img[3:8, 10 + gap:18 + gap] = 1.0
        img[8, 2:18 + gap] = 1.0  # the bar joining the two arms
        img[5, 2] = 3.0  # the peak
        bw = _component_mask(img, img, 1, True, _strel_disk(3))
        assert bw[img > 0].all()
        if closes:
            assert bw[3:8, 10:10 + gap].all()
        else:
            assert not bw[3:6, 12].any()  # the notch's middle column, out of reach of the bar's dilation
    assert _sphere(3).sum() == 123  # lattice points with x^2 + y^2 + z^2 <= 9


def test_2d_median_method_raises_like_the_reference():
    """psf_gen_new.m:71 calls medfilt, which MATLAB does not define."""
    import pytest
    with pytest.raises(ValueError, match="medfilt"):
        psf_gen_new(np.ones((20, 20), np.float32), 0.1, 0.1, 1.5, "median")

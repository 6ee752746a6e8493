"""Invariants that pin the back-projector oracle (oracle/bp_oracle.py).

No MATLAB is available to run omw_backprojector_generation.m, so its
toolbox steps stay "parity unpinned"; these checks fix what the mathematics
does: the back projector's DC gain is mask(DC) / (|OTF(0)|^2 + alpha) with
|OTF(0)| = 1, the apodised mask lies in [0, 1] and is 1 at DC, a symmetric
PSF gives a point-symmetric mask, every z-slice of the OTF mask is convex,
and the output is real and finite."""
import numpy as np
import pytest

from petakit5d_amd import synth


@pytest.fixture(scope="module")
def bo():
    import bp_oracle
    return bp_oracle


@pytest.mark.parametrize("shear", [0.0, 31.5])
def test_dc_gain_and_range(bo, shear):
    psf = synth.gaussian_psf((24, 24, 31), 1.0, 2.5, shear)
    b, omw, absc, mask, sk = bo.omw_backprojector_generation(psf, 0.005)
    assert np.isfinite(b).all()
    assert abs(float(b.astype(np.float64).sum()) - 1 / 1.005) < 1e-5
    assert abs(omw[0, 0, 0] - 1 / 1.005) < 1e-6
    # recovered apodisation weights ifftshift(mask) = OTF_bp_omw ./ (conj(OTF) ./ (|OTF|^2 + alpha)) lie in [0, 1]
    H = np.fft.fftn(np.roll(psf.astype(np.float64) / psf.sum(), [-(s // 2) for s in psf.shape], (0, 1, 2)))
    big = np.abs(H) > 1e-3
    wgt = (omw[big] * (np.abs(H[big]) ** 2 + 0.005) / np.conj(H[big])).real
    assert wgt.min() > -1e-4 and wgt.max() < 1 + 1e-4
    assert mask[12, 12, 15]  # DC (fftshift centre) is inside the mask
    assert absc[12, 12, 15] == absc.max()


def test_symmetric_psf_symmetric_mask(bo):
    psf = synth.gaussian_psf((25, 25, 31), 1.2, 3.0)  # odd sizes: point-symmetric about floor(p/2)
    _, _, _, mask, _ = bo.omw_backprojector_generation(psf, 0.005)
    np.testing.assert_array_equal(mask, mask[::-1, ::-1, ::-1])


def _runs_contiguous(v):
    nz = np.flatnonzero(v)
    return nz.size == 0 or nz[-1] - nz[0] + 1 == nz.size


def test_mask_slices_convex(bo):
    """The hull pass along x runs last (omw_backprojector_generation.m:107-109):
    every x-slice is a convex polygon rasterised on the grid, so each of its
    rows and columns is one contiguous run."""
    psf = synth.gaussian_psf((24, 24, 31), 1.0, 2.5, 20.0)
    _, _, _, mask, _ = bo.omw_backprojector_generation(psf, 0.005)
    for x in range(mask.shape[1]):
        s = mask[:, x, :]
        assert all(_runs_contiguous(r) for r in s) and all(_runs_contiguous(c) for c in s.T)


def test_hull_and_knn_primitives(bo):
    img = np.zeros((7, 7), bool)
    img[1, 1] = img[1, 5] = img[5, 3] = True
    h = bo.bwconvhull(img)
    assert h[1, 1] and h[1, 5] and h[5, 3] and h[3, 3] and not h[5, 1] and not h[0, 0]
    E = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]], float)
    Q = np.array([[0.4, 0, 0], [0.5, 0, 0]], float)
    i, d = bo.knn2(E, Q)
    assert i.tolist() == [[0, 1], [0, 1]]  # tie at 0.5 keeps the lower index first
    np.testing.assert_allclose(d, [[0.4, 0.6], [0.5, 0.5]])


def test_errors(bo):
    with pytest.raises(ValueError, match="nonzero PSF"):
        bo.omw_backprojector_generation(np.zeros((5, 5, 5)), 0.005)


# ---------------------------------------------------------------------------------------------------------------------
# BackProjector (WB) oracle: pinned by closed forms, not by retyping its formulas.
def test_wb_fwhm_tent_is_exact(bo):
    # a tent of half-width h peaking on a sample: linear interpolation between
    # samples is exact on its flanks, so FWHM = h (also for non-integer h)
    for h, c in [(6.0, 20.0), (9.0, 15.0), (5.5, 30.0), (7.25, 12.0)]:
        x = np.arange(1, 41, dtype=np.float64)
        y = np.maximum(0.0, 1.0 - np.abs(x - c) / h).astype(np.float32)
        assert abs(float(bo.wb_fwhm(y)) - h) < 1e-5


def test_wb_fwhm_step_and_no_crossing(bo):
    assert np.isnan(bo.wb_fwhm(np.array([0.0, 0.1, 0.6, 1.0, 1.0], np.float32)))  # no trailing edge
    with pytest.raises(IndexError):
        bo.wb_fwhm(np.array([0.0, 0.0, 0.0, 0.0], np.float32) + np.array([0.4, 0.45, 0.49, 0.45], np.float32))


def test_wb_gaussian_fwhm_close_to_analytic(bo):
    from petakit5d_amd import synth
    psf = synth.gaussian_psf((63, 63, 81), 2.0, 4.0, 0.0)
    fx, fy, fz = bo.wb_fwhm_psf(psf)
    assert abs(fx / (2.3548 * 2.0) - 1) < 0.03 and abs(fy / (2.3548 * 2.0) - 1) < 0.03
    assert abs(fz / (2.3548 * 4.0) - 1) < 0.03


def _psf(bo):
    from petakit5d_amd import synth
    return synth.gaussian_psf((33, 33, 41), 1.2, 3.0, 15.0)


def test_wb_traditional_is_flipped_psf(bo):
    p = _psf(bo)
    out, otf, _ = bo.wb_backprojector(p, "traditional")
    np.testing.assert_array_equal(out, p[::-1, ::-1, ::-1].astype(np.float64))
    # its OTF is the DFT of the centred flipped PSF: the DC term is the PSF sum
    assert abs(otf[0, 0, 0] - p.astype(np.float64).sum()) < 1e-9


def test_wb_wiener_identity(bo):
    p = _psf(bo)
    out, otf, info = bo.wb_backprojector(p, "wiener", 0.005)
    o = np.fft.fftn(np.fft.ifftshift(p[::-1, ::-1, ::-1].astype(np.float64)))
    on = o / np.abs(o).max()
    np.testing.assert_allclose(otf * (np.abs(on) ** 2 + 0.005), on, rtol=0, atol=1e-12)
    # PSF_bp is the centred inverse DFT: its sum is the DC gain
    assert abs(out.sum() - otf[0, 0, 0].real) < 1e-9


def test_wb_butterworth_gain_at_cutoff(bo):
    # odd size 33 (centre 17), iRes = 33/4 -> cutoff t = 4 pixels: voxel 21 on axis 1 has w = 1 exactly,
    # where 1/sqrt(1 + (1/beta^2 - 1) w^n) = beta; at the centre (DC) the mask is 1
    p = _psf(bo)
    beta = 0.3
    _, otf, info = bo.wb_backprojector(p, "butterworth", 0.005, beta, 10, 2, (33 / 4, 33 / 4, 41 / 5))
    m = np.fft.fftshift(otf.real)
    assert abs(info["cutoff"][0] - 4.0) < 1e-12
    assert abs(m[16, 16, 20] - 1.0) < 1e-7
    assert abs(m[20, 16, 20] - beta) < 1e-6
    assert np.all(np.diff(m[16:, 16, 20]) <= 0)  # monotone along the axis


def test_wb_wiener_butterworth_dc_and_alpha_one(bo):
    p = _psf(bo)
    _, otf_w, _ = bo.wb_backprojector(p, "wiener", 0.005)
    _, otf_wb, info = bo.wb_backprojector(p, "wiener-butterworth", 0.005, 1, 10)
    assert abs(otf_wb[0, 0, 0] - otf_w[0, 0, 0]) < 1e-12  # mask(DC) = 1
    assert info["beta"] == info["beta_fp"] and 0 < info["beta_fp"] < 1  # beta == 1 -> traditional cutoff gain
    _, _, info2 = bo.wb_backprojector(p, "wiener", 1)
    assert info2["alpha"] == info2["beta_fp"]


def test_wb_gaussian_back_projector_normalised(bo):
    p = _psf(bo)
    out, otf, info = bo.wb_backprojector(p, "gaussian")
    # the generated Gaussian uses the measured FWHMs: centred, unit integral (sampled), FWHMs reproduced
    assert np.unravel_index(np.argmax(out), out.shape) == (16, 16, 20)
    assert abs(out.sum() - 1) < 0.02
    f = bo.wb_fwhm_psf(out.astype(np.float32))
    assert all(abs(a / b - 1) < 0.05 for a, b in zip(f, info["FWHM"]))

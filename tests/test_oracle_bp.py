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

"""Pins the oracle (oracle/pk_oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible mistake fails it:
  * FFT: a textbook DFT matrix product and numpy's pocketfft (golden
    vectors), Parseval, round trip -- catches wrong signs, twiddles, strides;
  * decon_psf2otf: the direct DFT sum with the floor(p/2) origin, delta and
    Hermitian cases -- catches a wrong circshift or padding side;
  * RL / OMW: golden vectors from direct convolution / correlation sums
    (tests/golden/make_golden.py, no FFT) -- catches conv vs correlation
    (a missing conj), a dropped eps clamp, a wrong Biggs index or lambda;
    plus closed forms: delta-PSF fixed point, constant image, flux
    conservation, OMW(flip(PSF)) == RL for NUMIT <= 2;
  * epilogue: uint16 ties / saturation, bbox, damp, scale / offset;
  * findGoodFactorNumber: values worked out by hand from the rule.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dft_matrix(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def _naive_fftn(x):
    y = x.astype(np.complex128)
    for ax in range(3):
        y = np.moveaxis(np.tensordot(_dft_matrix(y.shape[ax]), np.moveaxis(y, ax, 0), axes=(1, 0)), 0, ax)
    return y


@pytest.mark.parametrize("shape", [(6, 10, 15), (8, 9, 25), (4, 4, 4), (7, 1, 3), (1, 1, 1), (2, 3, 5), (12, 5, 9),
                                   (11, 13, 2)])
def test_fft_vs_textbook_dft(oracle, shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = _naive_fftn(x)
    assert rel_l2(oracle.fftn(x, precision="f64"), ref) < 1e-13
    assert rel_l2(oracle.fftn(x, precision="f32"), ref) < 1e-6
    back = oracle.ifftn(oracle.fftn(x, precision="f64"), precision="f64")
    assert rel_l2(back, x) < 1e-13


def test_fft_golden_vectors(oracle):
    for name in ("fft_6x10x15.npz", "fft_real_8x9x25.npz"):
        g = np.load(os.path.join(GOLD, name))
        assert rel_l2(oracle.fftn(g["x"], precision="f64"), g["X"]) < 1e-13
        assert rel_l2(oracle.fftn(g["x"], precision="f32"), g["X"]) < 1e-6


@pytest.mark.parametrize("shape", [(64, 64, 101), (128, 128, 64), (101, 3, 2), (30, 27, 25)])
def test_fft_vs_numpy_and_parseval(oracle, shape):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape).astype(np.float32)
    X = oracle.fftn(x, precision="f32")
    ref = np.fft.fftn(x.astype(np.float64))
    assert rel_l2(X, ref) < 1e-6
    assert abs(np.sum(np.abs(X.astype(np.complex128)) ** 2) / x.size / np.sum(x.astype(np.float64) ** 2) - 1) < 1e-5


def test_psf2otf_direct_definition(oracle):
    rng = np.random.default_rng(3)
    psf = rng.random((4, 3, 5))
    od = (7, 6, 9)
    got = oracle.decon_psf2otf(psf, od, precision="f64")
    c = [s // 2 for s in psf.shape]
    k = np.stack(np.meshgrid(*[np.arange(n) for n in od], indexing="ij"), -1).reshape(-1, 3)
    ref = np.zeros(len(k), np.complex128)
    for o in np.ndindex(*psf.shape):
        ph = sum(k[:, d] * (o[d] - c[d]) / od[d] for d in range(3))
        ref += psf[o] * np.exp(-2j * np.pi * ph)
    assert rel_l2(got.reshape(-1, order="C"), ref.reshape(od).reshape(-1, order="C")) < 1e-12


def test_psf2otf_cases(oracle):
    d = np.zeros((5, 4, 7))
    d[2, 2, 3] = 1.0  # floor(p/2) origin -> all-ones OTF
    assert np.allclose(oracle.decon_psf2otf(d, (10, 8, 14), precision="f64"), 1.0, atol=1e-14)
    rng = np.random.default_rng(4)
    p = rng.random((3, 5, 4))
    H = oracle.decon_psf2otf(p, (9, 10, 8), precision="f64")
    Hm = np.conj(np.roll(H[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))  # H(-k)*
    assert rel_l2(H, Hm) < 1e-13
    assert not oracle.decon_psf2otf(np.zeros((3, 3, 3)), (6, 6, 6)).any()
    with pytest.raises(oracle.OracleError, match="outSizeIsSmallerThanPsfSize"):
        oracle.decon_psf2otf(p, (2, 10, 8))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_rl_golden_direct_sums(oracle, prec, tol):
    g = np.load(os.path.join(GOLD, "rl_8x6x5.npz"))
    for k in (1, 2, 3, 5):
        J, _, _ = oracle.decon_lucy_function(g["I"], g["psf"], k, Background=float(g["background"]),
                                             save16bit=False, precision=prec)
        assert rel_l2(J, g[f"J{k}"]) < tol, k


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_omw_golden_direct_sums(oracle, prec, tol):
    g = np.load(os.path.join(GOLD, "omw_7x6x4.npz"))
    for k in (2, 3):
        J, _ = oracle.decon_lucy_omw_function(g["I"], g["psf_f"], g["psf_b"], k, Background=float(g["background"]),
                                              save16bit=False, precision=prec)
        assert rel_l2(J, g[f"J{k}"]) < tol, k


def test_rl_delta_psf_fixed_point(oracle):
    rng = np.random.default_rng(5)
    I = rng.integers(0, 1000, (12, 10, 9)).astype(np.float32)
    psf = np.zeros((3, 3, 3), np.float32)
    psf[1, 1, 1] = 2.0
    for k in (1, 2, 6):
        J, _, _ = oracle.decon_lucy_function(I, psf, k, save16bit=False)
        assert rel_l2(J, I) < 1e-6


def test_rl_constant_image(oracle):
    I = np.full((16, 12, 10), 37.0, np.float32)
    psf = np.random.default_rng(6).random((5, 5, 3)).astype(np.float32)
    J, _, _ = oracle.decon_lucy_function(I, psf, 7, save16bit=False, precision="f64")
    assert np.allclose(J, 37.0, rtol=1e-12)
    J, _ = oracle.decon_lucy_omw_function(I, psf, psf, 3, save16bit=False, precision="f64")
    assert np.allclose(J, 37.0, rtol=1e-12)


def test_rl_flux_conservation(oracle):
    rng = np.random.default_rng(7)
    I = rng.integers(50, 500, (20, 16, 12)).astype(np.float64)
    psf = rng.random((5, 3, 5))
    for k in (1, 2):  # lambda = 0 for k <= 2: plain RL conserves flux
        J, _, _ = oracle.decon_lucy_function(I, psf, k, save16bit=False, precision="f64")
        assert abs(J.sum() / I.sum() - 1) < 1e-12


def test_omw_flipped_equals_rl(oracle):
    rng = np.random.default_rng(8)
    I = rng.integers(50, 500, (16, 12, 10)).astype(np.float32)
    psf = rng.random((5, 3, 7)).astype(np.float32)  # odd sizes: flip(p) has OTF conj(OTF(p))
    for k in (1, 2):
        R, _, _ = oracle.decon_lucy_function(I, psf, k, save16bit=False, precision="f64")
        O, _ = oracle.decon_lucy_omw_function(I, psf, psf[::-1, ::-1, ::-1], k, save16bit=False, precision="f64")
        assert rel_l2(O, R) < 1e-12


def test_biggs_lambda_changes_result(oracle):
    """From k = 3 on, the accelerated iterate differs from plain RL (the golden
    file stores the lambdas the literal formula gives)."""
    g = np.load(os.path.join(GOLD, "rl_8x6x5.npz"))
    assert g["lambdas"][0] == 0 and g["lambdas"][1] == 0 and g["lambdas"][2] > 0


def test_uint16_rounding(oracle):
    vals = [0.5, 1.5, 2.5, 0.49999997, 65534.5, 65535.4, 70000.0, -3.0, float("nan"), 2.4999999]
    exp = [1, 2, 3, 0, 65535, 65535, 65535, 0, 0, 2]
    assert [int(oracle.lib().pko_to_u16(v)) for v in vals] == exp
    assert list(oracle.to_uint16(np.array(vals))) == exp
    I = np.array([1, 3, 5, 0], np.uint16).reshape(4, 1, 1)
    J, _, _ = oracle.decon_lucy_function(I, np.ones((1, 1, 1), np.float32), 0, scaleFactor=0.5)
    assert list(J.ravel()) == [1, 2, 3, 0]


def test_bbox_damp_offset(oracle):
    rng = np.random.default_rng(9)
    I = rng.integers(0, 300, (10, 8, 6)).astype(np.uint16)
    one = np.ones((1, 1, 1), np.float32)
    J, _, _ = oracle.decon_lucy_function(I, one, 0, bbox=[2, 3, 1, 9, 8, 4], save16bit=True)
    np.testing.assert_array_equal(J, I[1:9, 2:8, 0:4])
    # offset with mask: zero raw voxels stay zero
    J, _, _ = oracle.decon_lucy_function(I, one, 0, deconOffset=5, save16bit=False)
    np.testing.assert_allclose(J, np.where(I != 0, I.astype(np.float32) + 5, 0))


def test_edge_erosion_full_rectangle(oracle):
    m = np.ones((8, 7, 6), bool)
    e = oracle.decon_mask_edge_erosion(m, 2)
    ref = np.zeros_like(m)
    ref[2:-2, 2:-2, 2:-2] = True
    np.testing.assert_array_equal(e, ref)


def test_edge_erosion_irregular(oracle):
    m = np.ones((9, 9, 7), bool)
    m[0, 0, :] = False  # irregular: faces cleared, then disk(1) in-plane and a 3-line in z
    e = oracle.decon_mask_edge_erosion(m, 2)
    ref = np.zeros_like(m)
    ref[2:-2, 2:-2, 2:-2] = True  # faces cleared -> cross erosion strips one more voxel
    np.testing.assert_array_equal(e, ref)


def test_strel_disk_small_radii_exact(oracle):
    """strel('disk', r) for r < 3 is the exact disk dx^2 + dy^2 <= r^2."""
    for r in (0, 1, 2):
        g = np.arange(-r, r + 1)
        np.testing.assert_array_equal(oracle.strel_disk(r), g[:, None] ** 2 + g[None, :] ** 2 <= r * r)


def test_strel_disk_radius5_documented_neighbourhood(oracle):
    """The MATLAB documentation shows strel('disk', 5) as a 9 x 9 neighbourhood
    (smaller than the radius-5 disk: the periodic-line approximation)."""
    want = np.array([[0, 0, 1, 1, 1, 1, 1, 0, 0],
                     [0, 1, 1, 1, 1, 1, 1, 1, 0],
                     [1, 1, 1, 1, 1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1, 1, 1, 1, 1],
                     [1, 1, 1, 1, 1, 1, 1, 1, 1],
                     [0, 1, 1, 1, 1, 1, 1, 1, 0],
                     [0, 0, 1, 1, 1, 1, 1, 0, 0]], bool)
    np.testing.assert_array_equal(oracle.strel_disk(5), want)


def _periodic_line_sum(r):
    """strel('disk', r), r >= 3, by brute force: dilate a point by each periodic
    line {i v : |i| <= floor(k / |v|)} (scipy), then by the two compensating
    lines."""
    from scipy import ndimage as ndi
    G = 4 * r + 3
    img = np.zeros((G, G), bool)
    img[G // 2, G // 2] = True
    t = np.pi / 8
    k = 2 * r / (1 / np.tan(t) + 1 / np.sin(t))
    for v in ((1, 0), (1, 1), (0, 1), (-1, 1)):
        rp = int(np.floor(k / np.hypot(*v)))
        se = np.zeros((2 * rp + 1,) * 2, bool) if rp else np.ones((1, 1), bool)
        for i in range(-rp, rp + 1):
            se[rp + i * v[0], rp + i * v[1]] = True
        img = ndi.binary_dilation(img, structure=se)
    rows = np.nonzero(img.any(axis=1))[0]
    e = rows.max() - G // 2
    ln = 2 * (r - e - 1) + 1
    if ln >= 3:
        img = ndi.binary_dilation(img, structure=np.ones((1, ln), bool))
        img = ndi.binary_dilation(img, structure=np.ones((ln, 1), bool))
    rr, cc = np.nonzero(img)
    R = max(np.abs(rr - G // 2).max(), np.abs(cc - G // 2).max())
    return img[G // 2 - R:G // 2 + R + 1, G // 2 - R:G // 2 + R + 1]


@pytest.mark.parametrize("r", [3, 4, 6, 7, 9, 12, 17, 25])
def test_strel_disk_is_the_periodic_line_decomposition(oracle, r):
    nh = oracle.strel_disk(r)
    np.testing.assert_array_equal(nh, _periodic_line_sum(r))
    assert r - 2 <= nh.shape[0] // 2 <= r  # an approximation from inside the radius-r extent
    np.testing.assert_array_equal(nh, nh[::-1, ::-1])
    np.testing.assert_array_equal(nh, nh.T)


@pytest.mark.parametrize("E", [4, 8])
def test_edge_erosion_irregular_large(oracle, E):
    """EdgeErosion 4 and 8 (the frame driver's default, XR_RLdeconFrame3D.m:41)
    on an irregular mask: faces cleared, every slice eroded by strel('disk',
    E - 1), then dim 3 by a (2E - 1)-long line (decon_mask_edge_erosion.m:
    17-27); out-of-volume neighbours count as true.  Checked against scipy's
    binary erosion with the same neighbourhoods."""
    from scipy import ndimage as ndi
    g = np.random.default_rng(E)
    m = np.ones((40, 36, 24), bool)
    for _ in range(6):  # holes of zeros (an irregular I ~= 0 mask)
        c = g.integers(0, [40, 36, 24])
        m[c[0]:c[0] + 5, c[1]:c[1] + 4, c[2]:c[2] + 3] = False
    e = oracle.decon_mask_edge_erosion(m, E)
    ref = m.copy()
    ref[[0, -1], :, :] = False
    ref[:, [0, -1], :] = False
    ref[:, :, [0, -1]] = False
    se = oracle.strel_disk(E - 1)
    ref = ndi.binary_erosion(ref, structure=se[:, :, None], border_value=1)
    ref = ndi.binary_erosion(ref, structure=np.ones((1, 1, 2 * E - 1), bool), border_value=1)
    np.testing.assert_array_equal(e, ref)


@pytest.mark.parametrize("n,d,odd7,mp,exp", [
    (101, 1, False, 7, 108), (257, 1, False, 7, 270), (448, 1, False, 7, 448), (448, 1, False, 5, 450),
    (7, 1, False, 7, 8), (14, 1, False, 7, 14), (21, 1, False, 7, 24), (21, 1, True, 7, 21),
    (101, -1, False, 7, 100), (1, 1, False, 7, 1), (101, 1, False, 5, 108), (129, 1, False, 5, 135),
])
def test_find_good_factor_number(oracle, n, d, odd7, mp, exp):
    assert oracle.findGoodFactorNumber(n, d, odd7, maxPrime=mp) == exp

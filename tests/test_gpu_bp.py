"""omw_backprojector_generation through the C-ABI (host toolbox steps + GPU
nearest-edge search) against the numpy/scipy oracle on the same PSFs."""
import numpy as np
import pytest

from conftest import rel_l2
from petakit5d_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bo():
    import bp_oracle
    return bp_oracle


@pytest.mark.parametrize("size,sig,shear,skewed", [
    ((24, 24, 31), (1.0, 2.5), 0.0, None),
    ((24, 24, 31), (1.0, 2.5), 31.5, None),
    ((32, 32, 41), (1.2, 3.0), 0.0, None),
    ((20, 22, 25), (1.0, 2.0), 15.0, True),
    ((24, 24, 31), (1.3, 3.2), 0.0, False),
])
def test_backprojector_vs_oracle(bo, size, sig, shear, skewed):
    import petakit5d_amd as pk
    print(f"[bp] size={size} sig={sig} shear={shear} skewed={skewed}", flush=True)
    psf = synth.gaussian_psf(size, sig[0], sig[1], shear)
    b, omw, absc, mask = pk.omw_backprojector_generation(psf, 0.005, skewed, OTFCumThresh=0.9,
                                                         hannWinBounds=[0.8, 1])
    rb, romw, rabsc, rmask, rsk = bo.omw_backprojector_generation(psf, 0.005, skewed, OTFCumThresh=0.9,
                                                                  hannWinBounds=(0.8, 1.0))
    assert b.shape == psf.shape and b.dtype == np.float32
    assert rel_l2(absc, rabsc) < 1e-6
    np.testing.assert_array_equal(np.asarray(mask, np.float64), np.asarray(rmask, np.float64))
    e = rel_l2(b, rb)
    print(f"  skewed={rsk} mask voxels={int(np.sum(rmask))} rel-L2(b_omw) {e:.2e}")
    assert e < 1e-5
    assert rel_l2(omw, romw) < 1e-5


def test_sheared_64x64x101_psf_backprojector_vs_oracle(bo):
    """The 64x64x101 sheared PSF of the 512x512x256 workload (413,696 voxels): OTF mask equal,
    b_omw and OTF_bp_omw within 1e-5 of the oracle, DC gain 1 / (1 + alpha)."""
    import petakit5d_amd as pk
    psf = synth.sheared_psf()
    print("[bp sheared psf] 64x64x101", flush=True)
    b, omw, absc, mask = pk.omw_backprojector_generation(psf, 0.005, None)
    rb, romw, rabsc, rmask, rsk = bo.omw_backprojector_generation(psf, 0.005, None)
    assert b.shape == (64, 64, 101) and np.isfinite(b).all()
    assert abs(float(b.astype(np.float64).sum()) - 1 / 1.005) < 1e-4
    np.testing.assert_array_equal(np.asarray(mask, np.float64), np.asarray(rmask, np.float64))
    e = rel_l2(b, rb)
    print(f"  skewed={rsk} mask voxels {float(np.sum(mask)):.0f}, rel-L2(b_omw) {e:.2e}")
    assert e < 1e-5 and rel_l2(omw, romw) < 1e-5


def test_backprojector_errors():
    import petakit5d_amd as pk
    with pytest.raises(pk.PkDeconError) as e:
        pk.omw_backprojector_generation(np.zeros((5, 5, 5), np.float32), 0.005)
    assert e.value.status == "PK_ERR_EMPTY_PSF"

"""RLdecon.m per-volume semantics without the device: the back projector cache
name (RLdecon.m:292) and the no-decon branch (nIter = 0 or an all-zero volume
returns the raw data, cropped to deconBbox; :415-420)."""
import numpy as np
import pytest

from petakit5d_amd.rldecon import RLdecon_volume, omw_bp_filename, psfgen_crop


def test_bp_cache_name_matches_reference_format():
    # sprintf('%s/%s_back_projector_alpha_%0.6f_otf_cum_thresh_%0.6f.tif', psfgenPath, psfFsn, 0.005, 0.9)
    fn = omw_bp_filename("/data/psfgen", "psf_488", 0.005, 0.9)
    assert fn == "/data/psfgen/psf_488_back_projector_alpha_0.005000_otf_cum_thresh_0.900000.tif"


def test_no_iterations_returns_raw_cropped():
    raw = np.asfortranarray(np.arange(6 * 5 * 4, dtype=np.uint16).reshape((6, 5, 4), order="F"))
    psf = np.ones((3, 3, 3), np.float32)
    out = RLdecon_volume(raw, psf, 0, RLMethod="simplified", deconBbox=[2, 1, 2, 5, 4, 3])
    np.testing.assert_array_equal(out, raw[1:5, 0:4, 1:3])
    zero = np.zeros((6, 5, 4), np.uint16)
    assert RLdecon_volume(zero, psf, 5, RLMethod="simplified") is zero


def test_unserved_method_rejected():
    with pytest.raises(ValueError):
        RLdecon_volume(np.ones((4, 4, 4), np.uint16), np.ones((3, 3, 3), np.float32), 2, RLMethod="cudagen")


def test_psfgen_crop_keeps_the_centre():
    # non-zero extent y 3..8 of 12 (0-based), x 5..6 of 9, z all of 7: crop min(first, n - 1 - last) - 1
    # per axis on both sides -> y 2, x 1, z 0
    p = np.zeros((12, 9, 7), np.float32)
    p[3:9, 5:7, :] = 1
    c = psfgen_crop(p)
    assert c.shape == (8, 7, 7)
    np.testing.assert_array_equal(c, p[2:10, 1:8, :])


def test_psfgen_dz_ratio():
    """dz_ratio = sind(|SkewAngle|) for stage scans, 1 for objective scans (RLdecon.m:101-106)."""
    import math
    from petakit5d_amd.rldecon import psfgen_dz_ratio
    assert psfgen_dz_ratio(32.45) == math.sin(32.45 * math.pi / 180)
    assert psfgen_dz_ratio(-32.45) == psfgen_dz_ratio(32.45)
    assert psfgen_dz_ratio(32.45, objectiveScan=True) == 1.0
    assert psfgen_dz_ratio(90) == 1.0 and psfgen_dz_ratio(0) == 0.0


def test_option_rules_of_the_reference():
    """interpMethod is validated with @isnumeric in RLdecon.m:26, so a method
    name is rejected; nTapering > 0 (edgetaper, :352-357) is not served;
    useGPU=false has no branch here.  All raise before any device work."""
    import petakit5d_amd as pk
    raw = np.ones((8, 8, 4), np.uint16)
    psf = np.ones((3, 3, 3), np.float32)
    with pytest.raises(ValueError, match="isnumeric"):
        pk.RLdecon_volume(raw, psf, 2, interpMethod="cubic")
    with pytest.raises(NotImplementedError, match="nTapering"):
        pk.RLdecon_volume(raw, psf, 2, nTapering=5)
    with pytest.raises(ValueError, match="useGPU"):
        pk.RLdecon_volume(raw, psf, 2, useGPU=False)

"""CPU checks of the C-ABI boundary: libpkdecon.so loads, exports every
symbol include/pkdecon.h declares, and its host-only logic behaves; plus the
Python front end's argument handling (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import petakit5d_amd as pk
from petakit5d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "pkdecon.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pk_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert _header_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in _header_functions():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_find_good_factor_number_matches_oracle(oracle):
    for n in range(1, 1500, 7):
        for mp, odd in ((7, False), (7, True), (5, False)):
            assert _lib.find_good_factor_number(n, 1, odd, mp) == oracle.findGoodFactorNumber(n, 1, odd, mp)
            assert _lib.find_good_factor_number(n, -1, odd, mp) == oracle.findGoodFactorNumber(n, -1, odd, mp)


def test_opts_default():
    o = _lib.PkOpts()
    _lib.load().pk_opts_default(ctypes.byref(o))
    assert (o.background, o.damp_factor, o.scale_factor, o.decon_offset, o.edge_erosion, o.save16bit,
            o.has_bbox) == (0.0, 1.0, 1.0, 0.0, 0, 1, 0)


def test_null_arguments_do_not_crash():
    L = _lib.load()
    assert L.pk_reset(None) == 1
    assert L.pk_decon_lucy(None, None, None, None, 1, None, None, None) == 1
    assert L.pk_last_error(None) == b"null context"
    assert L.pk_ctx_create(0, 0, None) == 1


def test_no_device_is_reported():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert _lib.load().pk_ctx_create(0, 0, ctypes.byref(h)) == 10  # PK_ERR_NO_DEVICE


def test_option_parsing():
    from petakit5d_amd.decon import _parse
    o = _parse({"background": 100, "SAVE16BIT": False, "dampfactor": 2})
    assert o["Background"] == 100 and o["save16bit"] is False and o["dampFactor"] == 2
    with pytest.raises(TypeError):
        _parse({"Backgroundd": 1})
    with pytest.raises(ValueError, match="useGPU"):
        pk.decon_lucy_function(np.zeros((4, 4, 4), np.uint16), np.ones((1, 1, 1)), 1, useGPU=False)


def test_crop_psf_to_data():
    p = np.arange(5 * 6 * 7, dtype=np.float32).reshape(5, 6, 7)
    c = pk.crop_psf_to_data(p, (3, 6, 4))
    # s = max(0, floor((p - im)/2)) + 1 (1-based), t = s + min(p, im) - 1
    np.testing.assert_array_equal(c, p[1:4, 0:6, 1:5])
    assert pk.crop_psf_to_data(p, (8, 8, 8)) is p


@pytest.mark.parametrize("bp_type", ["traditional", "gaussian", "butterworth", "wiener", "wiener-butterworth"])
def test_wb_backprojector_library_vs_oracle(bp_type):
    """BackProjector through the C-ABI (host computation, no device) against the numpy restatement."""
    import bp_oracle
    import petakit5d_amd as pk
    from petakit5d_amd import synth
    for shape, sig, shear, args in [((31, 31, 41), (1.2, 3.0), 0.0, (0.005, 1, 10, 1, (0, 0, 0))),
                                    ((24, 26, 33), (1.0, 2.5), 20.0, (1, 0.02, 8, 0, (0, 0, 0))),
                                    ((30, 28, 1), (1.3, 1.0), 0.0, (0.002, 1, 10, 2, (3.0, 3.0, 1.0)))]:
        if shape[2] == 1 and bp_type == "gaussian":
            continue  # FWHMz = 0 for a 2-D PSF: the reference's Gaussian is 0/0 (NaN) there, as here
        psf = synth.gaussian_psf(shape, sig[0], sig[1], shear) if shape[2] > 1 else \
            synth.gaussian_psf((30, 28, 3), sig[0], sig[1], 0.0)[:, :, 1:2]
        P, O, info = pk.BackProjector(psf, bp_type, *args, info=True)
        Pr, Or, ir = bp_oracle.wb_backprojector(psf, bp_type, *args)
        assert P.dtype == np.float64 and O.dtype == np.complex128 and P.shape == psf.shape
        assert info["FWHM"] == ir["FWHM"] and abs(info["beta_fp"] - ir["beta_fp"]) < 1e-12
        for a, b in ((P, Pr), (O, Or)):
            assert np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300) < 1e-6


@pytest.mark.parametrize("bp_type", ["traditional", "wiener", "butterworth"])
def test_wb_backprojector_step_like_profile(bp_type):
    """A z profile that never falls back below half maximum has no FWHM
    (fwhm returns NaN, BackProjector.m:474-509): the cutoff sample indices
    are then max(round(NaN), 1) = 1 and min(round(NaN), Sz) = Sz, MATLAB's max
    and min skipping NaN (BackProjector.m:127-129).  Library and oracle agree,
    NaN for NaN where the Butterworth mask carries the NaN cutoff."""
    import bp_oracle
    import petakit5d_amd as pk
    from petakit5d_amd import synth
    pz = np.array([0.1, 0.3, 0.7, 1.0, 0.9, 0.8, 0.7], np.float32)
    psf = (synth.gaussian_psf((24, 26, 3), 1.1, 1.0, 0.0)[:, :, 1:2] * pz[None, None, :]).astype(np.float32)
    P, O, info = pk.BackProjector(psf, bp_type, 0.005, 1, 10, 1, info=True)
    Pr, Or, ir = bp_oracle.wb_backprojector(psf, bp_type, 0.005, 1, 10, 1)
    assert np.isnan(ir["FWHM"][2]) and np.isnan(info["FWHM"][2])
    assert np.isfinite(ir["beta_fp"]) and abs(info["beta_fp"] - ir["beta_fp"]) < 1e-12
    for a, b in ((P, Pr), (O, Or)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        f = ~np.isnan(b)
        if f.any():
            assert np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(b[f]), 1e-300) < 1e-6
    if bp_type == "butterworth":
        assert np.isnan(P).all()


def test_wb_backprojector_errors():
    import petakit5d_amd as pk
    p = np.ones((9, 9, 9), np.float32)
    with pytest.raises(ValueError):
        pk.BackProjector(p, "nope")
    with pytest.raises(pk.PkDeconError) as e:
        pk.BackProjector(p, "wiener", 0.005, 1, 10, 3)
    assert e.value.status == "PK_ERR_INVALID_ARG"
    with pytest.raises(pk.PkDeconError):
        pk.BackProjector(np.ones((1, 9, 9), np.float32), "wiener")

"""deskewRotateFrame3D on the device (pk_deskew_rotate_frame3d) against the
oracle (oracle/deskew_oracle.py: skewed-space interpolation of
skewed_space_interp_defined_stepsize.m, then the row warp of
volume_deskew_rotate_warp_mex.cpp, or imwarp's trilinear form for objective
scans): forward and reversed stage scans with and without the finer-step
interpolation, uint16 input with uint16 output, a bbox, an objective scan,
and RLdecon's default DSRCombined path."""
import numpy as np
import pytest

import deskew_oracle as D

pytestmark = pytest.mark.gpu


def _vol(shape, seed, u16=False):
    g = np.random.default_rng(seed)
    v = g.random(shape).astype(np.float32) * 1000
    return np.asfortranarray(v.astype(np.uint16) if u16 else v)


@pytest.mark.parametrize("dz", [0.3, 0.2, 0.5])   # dx = 2.34 (step 1/2), 1.56 (none), 3.91 (step 1/2)
@pytest.mark.parametrize("reverse", [False, True])
def test_dsr_vs_oracle(dz, reverse):
    import petakit5d_amd as pk
    print(f"[dsr] dz={dz} reverse={reverse}", flush=True)
    v = _vol((48, 120, 40), 1)
    got = pk.deskewRotateFrame3D(v, 32.45, dz, 0.108, reverse=reverse)
    ref = D.deskew_rotate(v, 32.45, dz, 0.108, reverse=reverse)
    assert got.shape == ref.shape and got.dtype == np.float32
    err = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
    print(f"  out {got.shape} max rel err {err:.2e}")
    assert err < 1e-5


def test_dsr_uint16_bbox_and_objective_scan():
    import petakit5d_amd as pk
    v16 = _vol((32, 96, 36), 2, u16=True)
    ref = D.deskew_rotate(v16.astype(np.float32), 32.45, 0.3, 0.108)
    got = pk.deskewRotateFrame3D(v16, 32.45, 0.3, 0.108, save16bit=True)
    assert got.dtype == np.uint16
    d = np.abs(got.astype(np.int64) - np.clip(np.floor(ref.astype(np.float64) + 0.5), 0, 65535).astype(np.int64))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    box = [3, 5, 2, 30, ref.shape[1] - 4, ref.shape[2] - 3]
    gb = pk.deskewRotateFrame3D(v16.astype(np.float32), 32.45, 0.3, 0.108, bbox=box)
    np.testing.assert_allclose(gb, ref[box[0] - 1:box[3], box[1] - 1:box[4], box[2] - 1:box[5]], rtol=1e-5, atol=1e-2)
    vo = _vol((24, 40, 20), 3)
    go = pk.deskewRotateFrame3D(vo, 20.0, 0.3, 0.1, objectiveScan=True)
    ro = D.deskew_rotate(vo, 20.0, 0.3, 0.1, objectiveScan=True)
    assert go.shape == ro.shape
    assert np.abs(go - ro).max() / np.abs(ro).max() < 1e-5


def test_rldecon_default_dsr_combined():
    """RLdecon's default (Deskew, Rotate, DSRCombined true) returns
    deskewRotateFrame3D of the deconvolved volume (RLdecon.m:488-493), uint16
    with save16bit; a device tensor input stays on the device."""
    import torch
    import petakit5d_amd as pk
    from petakit5d_amd import synth
    psf = synth.gaussian_psf((15, 15, 21), 1.2, 3.0, 20.0)
    g0, g1 = synth.rng(71, 0), synth.rng(71, 1)
    img = synth._fft_convolve_circular(synth.point_emitters((64, 96, 40), 40, (500, 3000), g0), psf) + 100
    raw = synth.poisson_u16(img, g1)
    J, ds, dsr = pk.RLdecon_volume(raw, psf, 3, Background=100, Deskew=True, Rotate=True, SkewAngle=-32.45, dz=0.3,
                                   xyPixelSize=0.108)
    assert ds is None and dsr.dtype == np.uint16
    want = pk.deskewRotateFrame3D(J, 32.45, 0.3, 0.108, save16bit=True)
    np.testing.assert_array_equal(dsr, want)
    rt = torch.from_numpy(np.ascontiguousarray(raw.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    _, _, dsr_t = pk.RLdecon_volume(rt, psf, 3, Background=100, Deskew=True, Rotate=True, dz=0.3, xyPixelSize=0.108)
    assert dsr_t.is_cuda and dsr_t.dtype == torch.uint16
    np.testing.assert_array_equal(dsr_t.cpu().numpy(), dsr)

"""deskewRotateFrame3D with a resampleFactor on the device
(pk_deskew_rotate_frame3d_resample) against the oracle's restatement of the
reference's imwarp form (oracle/deskew_oracle.py: deskew_rotate_general;
deskewRotateFrame3D.m:72-73, 98-110, 183-210): linear and cubic, forward and
reversed, with and without skewed-space interpolation, uint16 with
save16bit, a bbox window, and ones falling back to the fast path bit for
bit; gpuProcess = true taking the same imwarp form without resampling."""
import numpy as np
import pytest

import deskew_oracle as D

pytestmark = pytest.mark.gpu


def _smooth(shape, seed):
    from scipy import ndimage as ndi
    g = np.random.default_rng(seed)
    v = ndi.gaussian_filter(g.random(shape) * 4000, 1.5) + g.random(shape) * 50
    return np.asfortranarray(v.astype(np.float32))


def _relmax(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("rs", [(1.0, 2.0, 1.0), (1.5, 0.5, 2.0)])
@pytest.mark.parametrize("dz,reverse", [(0.5, False), (0.5, True), (0.2, False)])
def test_dsr_resample_vs_oracle(interp, rs, dz, reverse):
    """dz 0.5: dx = 3.91 > 2, skewed-space interpolation first; dz 0.2: none."""
    import petakit5d_amd as pk
    print(f"[dsr resample] {interp} rs={rs} dz={dz} reverse={reverse}", flush=True)
    v = _smooth((12, 60, 24), 21)
    got = pk.deskewRotateFrame3D(v, 32.45, dz, 0.108, reverse=reverse, resampleFactor=rs, interpMethod=interp)
    ref = D.deskew_rotate_general(v, 32.45, dz, 0.108, reverse=reverse, resample=rs, interp=interp)
    assert got.shape == ref.shape and got.dtype == np.float32
    e = _relmax(got, ref)
    print(f"  out {got.shape} max rel err {e:.2e}", flush=True)
    assert e < 1e-5


def test_dsr_resample_uint16_bbox_and_ones():
    import petakit5d_amd as pk
    v16 = np.asfortranarray(_smooth((10, 64, 24), 22).astype(np.uint16))
    # uint16 with save16bit and skewed-space interpolation: the interpolated
    # slices are uint16 values, the result uint16
    got = pk.deskewRotateFrame3D(v16, 32.45, 0.5, 0.108, save16bit=True, resampleFactor=(1.0, 1.0, 0.5))
    ref = D.deskew_rotate_general(v16, 32.45, 0.5, 0.108, resample=(1.0, 1.0, 0.5), round_interp=True)
    assert got.dtype == np.uint16 and got.shape == ref.shape
    d = np.abs(got.astype(np.int64) - np.clip(np.floor(ref.astype(np.float64) + 0.5), 0, 65535).astype(np.int64))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    # a bbox is a window of the resampled output
    vf = _smooth((10, 64, 24), 23)
    rs = (2.0, 1.0, 1.5)
    full = D.deskew_rotate_general(vf, 32.45, 0.3, 0.108, resample=rs)
    box = [2, 3, 2, 8, full.shape[1] - 4, full.shape[2] - 1]
    gb = pk.deskewRotateFrame3D(vf, 32.45, 0.3, 0.108, bbox=box, resampleFactor=rs)
    rb = full[box[0] - 1:box[3], box[1] - 1:box[4], box[2] - 1:box[5]]
    assert gb.shape == rb.shape and _relmax(gb, rb) < 1e-5
    # ones: the fast path, bit for bit
    a = pk.deskewRotateFrame3D(vf, 32.45, 0.5, 0.108)
    b = pk.deskewRotateFrame3D(vf, 32.45, 0.5, 0.108, resampleFactor=[1, 1, 1])
    np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        pk.deskewRotateFrame3D(vf, 32.45, 0.5, 0.108, resampleFactor=[2, 2])


@pytest.mark.parametrize("dz,reverse", [(0.5, False), (0.5, True), (0.2, False)])
def test_dsr_gpu_process_takes_the_imwarp_form(dz, reverse):
    """gpuProcess = true leaves the fast path in the reference
    (deskewRotateFrame3D.m:73): the skewed-space interpolation when the step
    is large, then imwarp's trilinear form over outSize -- which differs from
    the fast path's row warp by the fast path's offset."""
    import petakit5d_amd as pk
    print(f"[dsr gpuProcess] dz={dz} reverse={reverse}", flush=True)
    v = _smooth((12, 60, 24), 24)
    got = pk.deskewRotateFrame3D(v, 32.45, dz, 0.108, reverse=reverse, gpuProcess=True)
    ref = D.deskew_rotate_general(v, 32.45, dz, 0.108, reverse=reverse)
    assert got.shape == ref.shape and got.dtype == np.float32
    assert _relmax(got, ref) < 1e-5
    v16 = np.asfortranarray(v.astype(np.uint16))
    g16 = pk.deskewRotateFrame3D(v16, 32.45, dz, 0.108, reverse=reverse, gpuProcess=True)
    assert g16.dtype == (np.float32 if dz == 0.5 else np.uint16)  # imwarp keeps the class of what it resamples

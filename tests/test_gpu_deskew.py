"""deskewFrame3D on the device (pk_deskew_frame3d) against the oracle: forward,
reverse and cropped geometries at the reference's typical angles, bitwise for
integer shifts, within float32 rounding of the float64 oracle otherwise; a
device tensor in, a device tensor out."""
import numpy as np
import pytest

import deskew_oracle as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("angle,dz,xy", [(32.45, 0.3, 0.108), (31.5, 0.5, 0.1), (60.0, 0.4, 0.1)])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("crop", [False, True])
@pytest.mark.parametrize("ny", [67, 64])  # scalar and float4 column paths
def test_deskew_vs_oracle(angle, dz, xy, reverse, crop, ny):
    from petakit5d_amd import deskewFrame3D
    print(f"[deskew] angle={angle} dz={dz} xy={xy} reverse={reverse} crop={crop} ny={ny}", flush=True)
    g = np.random.default_rng(2)
    vol = (g.random((ny, 48, 33)) * 1000).astype(np.float32)
    got = deskewFrame3D(vol, angle, dz, xy, reverse, Crop=crop)
    ref = D.deskew(vol, angle, dz, xy, reverse, crop)
    assert got.shape == ref.shape and got.dtype == np.float32
    # a + f (b - a) in float32: within a few ulps of the largest operand
    np.testing.assert_allclose(got, ref, rtol=0, atol=4e-7 * 1000 * 4)


def test_integer_shift_bitwise_and_uint16_input():
    from petakit5d_amd import deskewFrame3D
    g = np.random.default_rng(3)
    for ny in (130, 128):
        vol = g.integers(0, 65535, size=(ny, 40, 12)).astype(np.uint16)
        got = deskewFrame3D(vol, 0.0, 0.2, 0.1)
        np.testing.assert_array_equal(got, D.deskew(vol, 0.0, 0.2, 0.1).astype(np.float32))


def test_device_tensor_round_trip():
    import torch
    from petakit5d_amd import deskewFrame3D
    g = np.random.default_rng(4)
    vol = g.random((64, 50, 20)).astype(np.float32)
    t = torch.from_numpy(vol.transpose(2, 1, 0).copy()).cuda().permute(2, 1, 0)
    out = deskewFrame3D(t, 32.45, 0.3, 0.108, Crop=True)
    assert out.is_cuda and out.stride()[0] == 1
    host = out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0)
    np.testing.assert_array_equal(host, deskewFrame3D(vol, 32.45, 0.3, 0.108, Crop=True))

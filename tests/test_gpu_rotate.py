"""rotateFrame3D on the device (pk_affine_warp3d) against scipy's
map_coordinates (linear, zero outside, 'grid-constant') on the same transform:
a deskewed stage-scan volume with the automatic crop, reversed, objective
scan, resampled; a device tensor in, a device tensor out."""
import numpy as np
import pytest

from test_rotate import _warp_ref

pytestmark = pytest.mark.gpu


def _deskewed(seed=0):
    import petakit5d_amd as pk
    g = np.random.default_rng(seed)
    vol = (g.random((24, 40, 18)) * 1000).astype(np.float32)
    return pk.deskewFrame3D(vol, 32.45, 0.3, 0.108)  # zeros outside the sheared parallelogram


@pytest.mark.parametrize("kw", [dict(), dict(reverse=True), dict(Crop=False), dict(ObjectiveScan=True),
                                dict(resample=[1.5, 1.0, 2.0]), dict(outSize=[24, 50, 40])])
def test_rotate_vs_map_coordinates(kw):
    import petakit5d_amd as pk
    from petakit5d_amd.rotate import rotate_transform
    print(f"[rotate] {kw}", flush=True)
    ds = _deskewed()
    zx = np.sin(32.45 * np.pi / 180) * 0.3 / 0.108
    got = pk.rotateFrame3D(ds, 32.45, zx, **kw)
    T, od = rotate_transform(ds.shape, 32.45, zx, proj=None if kw.get("ObjectiveScan") else ds.max(axis=0), **kw)
    assert got.shape == od
    ref = _warp_ref(ds, T, od)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1000 * 2e-6)


def test_device_tensor():
    import torch
    import petakit5d_amd as pk
    ds = _deskewed(1)
    t = torch.from_numpy(ds.transpose(2, 1, 0).copy()).cuda().permute(2, 1, 0)
    out = pk.rotateFrame3D(t, 32.45, 1.5)
    assert out.is_cuda
    np.testing.assert_array_equal(out.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0), pk.rotateFrame3D(ds, 32.45, 1.5))

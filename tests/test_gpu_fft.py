"""GPU transform parity: libpkdecon's R2C / C2R passes against numpy's FFT
(pocketfft, an independent implementation) on column-major volumes."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [
    (8, 6, 5), (16, 16, 16), (9, 10, 15), (64, 64, 32), (128, 128, 64), (30, 27, 25), (2, 3, 5),
    (1, 4, 4), (4, 4, 1), (6, 1, 1), (512, 8, 4), (6, 250, 8), (10, 8, 243), (2048, 2, 2), (4, 2048, 2),
    (4, 4, 1024), (270, 9, 4), (375, 4, 2), (96, 100, 64), (486, 6, 10), (54, 64, 45),
]


def _ref_half(x):
    X = np.fft.rfft(x.astype(np.float64), axis=0)
    return np.fft.fft(np.fft.fft(X, axis=1), axis=2)


def _col_major_cuda(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda()
    return t.permute(2, 1, 0)


@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_c2r_vs_numpy(shape):
    import petakit5d_amd as pk
    print(f"[fft] shape={shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape).astype(np.float32)
    spec = pk.fft_r2c(_col_major_cuda(x))
    got = spec.cpu().numpy()
    ref = _ref_half(x)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < 2e-6
    back = pk.fft_c2r(spec, shape).cpu().numpy() / np.prod(shape)
    assert rel_l2(back, x) < 2e-6

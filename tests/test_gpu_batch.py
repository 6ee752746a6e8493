"""Batched decon (one launch per pass over all chunks) against one
single-volume call per chunk: results must be bitwise identical, including
chunks padded to 2/3/5-smooth sizes and chunks smaller than the PSF (PSF
centre-cropped per chunk as RLdecon.m:264-270 does)."""
import numpy as np
import pytest

from petakit5d_amd import synth

pytestmark = pytest.mark.gpu


def _chunks(sizes, seed=4):
    out = []
    for i, s in enumerate(sizes):
        g = synth.rng(seed * 100 + i, 0)
        out.append(g.integers(80, 400, size=s).astype(np.uint16))
    return out


SIZES = [(64, 48, 40), (50, 45, 32), (72, 64, 27), (33, 40, 20), (96, 80, 36), (64, 64, 30), (16, 1024, 30)]


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 5)])
def test_batch_matches_single_calls(method, numit):
    import petakit5d_amd as pk
    print(f"[batch] method={method} numit={numit}", flush=True)
    psf = synth.gaussian_psf((15, 15, 25), 1.2, 3.0, 15.0)  # larger than some chunks in z
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    chunks = _chunks(SIZES)
    outs = pk.decon_batch(chunks, psf, psf_b, numit, method=method, Background=100, save16bit=False)
    for c, J in zip(chunks, outs):
        pf = pk.crop_psf_to_data(psf, c.shape)
        pb = pk.crop_psf_to_data(psf_b, c.shape)
        if method == "omw":
            R, _ = pk.decon_lucy_omw_function(c, pf, pb, numit, Background=100, save16bit=False, pad=True)
        else:
            R, _, _ = pk.decon_lucy_function(c, pf, numit, Background=100, save16bit=False, pad=True)
        assert J.shape == c.shape
        np.testing.assert_array_equal(J, R)


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 3)])
def test_many_chunks_match_single_calls(method, numit):
    """More chunks than chunk-set streams: each stream runs several chunks
    and several block-shape groups per pass; every result stays bitwise equal
    to its single call."""
    import petakit5d_amd as pk
    g = synth.rng(77, 0)
    sizes = [(int(g.integers(20, 72)), int(g.integers(20, 72)), int(g.integers(10, 30))) for _ in range(20)]
    print(f"[batch many] method={method} numit={numit} sizes={sizes}", flush=True)
    psf = synth.gaussian_psf((9, 9, 11), 1.2, 3.0, 10.0)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    chunks = _chunks(sizes, seed=9)
    outs = pk.decon_batch(chunks, psf, psf_b, numit, method=method, Background=100, save16bit=False)
    for c, J in zip(chunks, outs):
        pf = pk.crop_psf_to_data(psf, c.shape)
        pb = pk.crop_psf_to_data(psf_b, c.shape)
        if method == "omw":
            R, _ = pk.decon_lucy_omw_function(c, pf, pb, numit, Background=100, save16bit=False, pad=True)
        else:
            R, _, _ = pk.decon_lucy_function(c, pf, numit, Background=100, save16bit=False, pad=True)
        np.testing.assert_array_equal(J, R)


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 3)])
def test_odd_row_chunks_match_single_calls(method, numit):
    """Chunks whose rows are odd after padding (43 -> 45, 131 -> 135, 26 -> 27,
    221 -> 225) or odd and smooth (75): the batched odd-row kernels read the
    raw chunk with its padding in-kernel, the single calls get a host-padded
    volume; results stay bitwise equal, uint16 output included."""
    import petakit5d_amd as pk
    sizes = [(43, 30, 20), (131, 12, 10), (26, 20, 12), (221, 9, 8), (75, 16, 12)]
    print(f"[batch odd] method={method} numit={numit}", flush=True)
    psf = synth.gaussian_psf((9, 9, 11), 1.2, 3.0, 10.0)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    chunks = _chunks(sizes, seed=11)
    for s16 in (False, True):
        outs = pk.decon_batch(chunks, psf, psf_b, numit, method=method, Background=100, save16bit=s16)
        for c, J in zip(chunks, outs):
            pf = pk.crop_psf_to_data(psf, c.shape)
            pb = pk.crop_psf_to_data(psf_b, c.shape)
            if method == "omw":
                R, _ = pk.decon_lucy_omw_function(c, pf, pb, numit, Background=100, save16bit=s16, pad=True)
            else:
                R, _, _ = pk.decon_lucy_function(c, pf, numit, Background=100, save16bit=s16, pad=True)
            np.testing.assert_array_equal(J, R)


def test_batch_device_chunks_uint16():
    import torch
    import petakit5d_amd as pk
    psf = synth.gaussian_psf((9, 9, 11), 1.2, 3.0)
    chunks = _chunks([(64, 48, 40), (48, 40, 30)], seed=9)
    dev = [torch.from_numpy(np.ascontiguousarray(c.transpose(2, 1, 0))).cuda().permute(2, 1, 0) for c in chunks]
    outs = pk.decon_batch(dev, psf, psf[::-1, ::-1, ::-1].copy(), 2, method="omw", Background=100)
    host = pk.decon_batch(chunks, psf, psf[::-1, ::-1, ::-1].copy(), 2, method="omw", Background=100)
    for d, h in zip(outs, host):
        assert d.is_cuda and d.dtype == torch.uint16
        np.testing.assert_array_equal(d.cpu().numpy(), h)


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 3), ("omw", 0)])
@pytest.mark.parametrize("kw", [
    dict(save16bit=True),
    dict(save16bit=False, scaleFactor=2.0, deconOffset=5.0),
    dict(save16bit=True, dampFactor=1.5, scaleFactor=0.5),
], ids=lambda k: "-".join(f"{a}={b}" for a, b in k.items()))
@pytest.mark.parametrize("dtype", ["uint16", "single"])
def test_batch_epilogue_options_match_single(method, numit, kw, dtype):
    """The batched last update writes each chunk's output through the epilogue
    and the prep pass reads each chunk in its own (unpadded) layout; every
    epilogue option must give the bits of one single-volume call per chunk."""
    import petakit5d_amd as pk
    print(f"[batch epilogue] {method} numit={numit} {kw} {dtype}", flush=True)
    psf = synth.gaussian_psf((9, 9, 11), 1.2, 3.0)
    psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
    # padded to 64x48x40 (no padding), 50x45x36, 48x40x27 and 45x40x24 (odd rows, padded)
    chunks = _chunks([(64, 48, 40), (50, 45, 33), (47, 40, 27), (43, 40, 21)], seed=12)
    for c in chunks:
        c[:2, :, :] = 0  # an I ~= 0 mask inside each chunk
    if dtype == "single":
        chunks = [c.astype(np.float32) for c in chunks]
    outs = pk.decon_batch(chunks, psf, psf_b, numit, method=method, Background=100, **kw)
    for c, J in zip(chunks, outs):
        if method == "omw":
            R, _ = pk.decon_lucy_omw_function(c, psf, psf_b, numit, Background=100, pad=True, **kw)
        else:
            R, _, _ = pk.decon_lucy_function(c, psf, numit, Background=100, pad=True, **kw)
        assert J.shape == c.shape and J.dtype == R.dtype
        np.testing.assert_array_equal(J, R)

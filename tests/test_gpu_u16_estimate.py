"""uint16 input keeps I = sat(raw - uint16(Background)) as uint16 and reads it
so in the ratio and OMW's first update; single input computes max(raw -
Background, 0) in float (decon_lucy_function.m:60-66,
decon_lucy_omw_function.m:55-60).  For integer data and an integer background
the two estimates hold the same numbers, so a uint16 volume and its float32
copy must deconvolve to bitwise the same float output: single calls (even
and odd rows, 1024-thread rows, a runtime-plan row length) and batches.  The
inputs include voxels below the background (I = 0 in both)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(shape, psf_shape, seed=5):
    from petakit5d_amd import synth
    psf = synth.gaussian_psf(psf_shape, 1.2, 3.0, 15.0)
    g0, g1 = synth.rng(seed, 0), synth.rng(seed, 1)
    obj = synth.point_emitters(shape, 40, (500, 3000), g0)
    img = synth._fft_convolve_circular(obj, psf) + 100.0
    return synth.poisson_u16(img, g1), psf


def _psf_b(psf):
    rng = np.random.default_rng(11)
    return (psf[::-1, ::-1, ::-1] * (1 + 0.3 * rng.standard_normal(psf.shape))).astype(np.float32)


@pytest.mark.parametrize("shape,psf_shape", [
    ((64, 64, 32), (15, 15, 21)),
    ((45, 30, 27), (9, 9, 11)),      # odd rows
    ((1024, 12, 10), (9, 9, 7)),     # 512-point rows
    ((96, 30, 20), (9, 9, 7)),       # 48-point rows (runtime-plan kernel: float I)
])
@pytest.mark.parametrize("method,numit", [("omw", 2), ("omw", 3), ("rl", 1), ("rl", 4)])
def test_uint16_matches_float_input_bitwise(shape, psf_shape, method, numit):
    import petakit5d_amd as pk
    print(f"[u16 estimate] shape={shape} method={method} numit={numit}", flush=True)
    I, psf = _case(shape, psf_shape)
    assert (I < 100).any() and (I > 100).any()
    If = I.astype(np.float32)
    if method == "omw":
        psf_b = _psf_b(psf)
        J16, _ = pk.decon_lucy_omw_function(I, psf, psf_b, numit, Background=100, save16bit=False)
        Jf, _ = pk.decon_lucy_omw_function(If, psf, psf_b, numit, Background=100, save16bit=False)
    else:
        J16, _, _ = pk.decon_lucy_function(I, psf, numit, Background=100, save16bit=False)
        Jf, _, _ = pk.decon_lucy_function(If, psf, numit, Background=100, save16bit=False)
    assert J16.dtype == np.float32 and np.isfinite(J16).all() and J16.max() > 0
    assert np.array_equal(J16.view(np.uint32), Jf.view(np.uint32))


@pytest.mark.parametrize("method,numit", [("omw", 2), ("rl", 3)])
def test_batch_uint16_matches_float_input_bitwise(method, numit):
    import petakit5d_amd as pk
    shapes = [(64, 64, 32), (45, 30, 27), (50, 36, 20)]
    cases = [_case(s, (9, 9, 11), seed=5 + i) for i, s in enumerate(shapes)]
    psf = cases[0][1]
    psf_b = _psf_b(psf)
    u16 = [c[0] for c in cases]
    f32 = [c.astype(np.float32) for c in u16]
    print(f"[u16 estimate batch] method={method} numit={numit}", flush=True)
    o16 = pk.decon_batch(u16, psf, psf_b, numit, method=method, Background=100, save16bit=False)
    of = pk.decon_batch(f32, psf, psf_b, numit, method=method, Background=100, save16bit=False)
    for a, b in zip(o16, of):
        assert a.dtype == np.float32 and np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))

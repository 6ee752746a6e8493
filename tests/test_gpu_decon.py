"""GPU decon parity through the C-ABI: decon_lucy_function (Biggs RL) and
decon_lucy_omw_function against the oracle's CPU single-precision path
(decon_lucy_function.m / decon_lucy_omw_function.m, useGPU=false) on the same
seeded inputs.  Bar: relative L2 <= 1e-4 on the float output; the FP64
oracle's error is printed beside it."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _case(shape, psf_shape, sig=(1.2, 3.0), n=60, seed=3, shear=0.0):
    from petakit5d_amd import synth
    psf = synth.gaussian_psf(psf_shape, sig[0], sig[1], shear)
    g0, g1 = synth.rng(seed, 0), synth.rng(seed, 1)
    obj = synth.point_emitters(shape, n, (500, 3000), g0)
    img = synth._fft_convolve_circular(obj, psf) + 100.0
    return synth.poisson_u16(img, g1), psf


def _flip(p):
    return np.ascontiguousarray(p[::-1, ::-1, ::-1])


@pytest.mark.parametrize("shape,psf_shape,numit", [
    ((64, 64, 32), (15, 15, 21), 15),
    ((45, 30, 27), (9, 9, 11), 7),
    ((50, 36, 1), (7, 7, 1), 5),
    ((256, 20, 12), (9, 9, 7), 6),   # 128-point rows
    ((2048, 6, 8), (7, 5, 5), 4),    # 1024-point rows (a row spans two wavefronts)
])
def test_rl_biggs_vs_oracle(oracle, shape, psf_shape, numit):
    import petakit5d_amd as pk
    print(f"[rl] shape={shape} psf={psf_shape} numit={numit}", flush=True)
    I, psf = _case(shape, psf_shape)
    J, err, k = pk.decon_lucy_function(I, psf, numit, Background=100, save16bit=False)
    R32, _, _ = oracle.decon_lucy_function(I, psf, numit, Background=100, save16bit=False)
    R64, _, _ = oracle.decon_lucy_function(I, psf, numit, Background=100, save16bit=False, precision="f64")
    e32, e64 = rel_l2(J, R32), rel_l2(J, R64)
    print(f"  rel-L2 vs CPU FP32 {e32:.3e}  vs FP64 {e64:.3e}  (CPU FP32 vs FP64 {rel_l2(R32, R64):.3e})")
    assert J.shape == shape and J.dtype == np.float32 and k == numit
    assert err.shape == (numit, 4) and not err.any()
    assert e32 <= TOL


@pytest.mark.parametrize("shape,psf_shape,numit", [
    ((64, 64, 32), (15, 15, 21), 2),
    ((48, 40, 30), (11, 13, 9), 3),
    ((64, 40, 30), (11, 13, 9), 0),   # no iteration: the epilogue of I
    ((64, 40, 30), (11, 13, 9), 1),
    ((1024, 12, 10), (9, 9, 7), 2),   # 512-point rows
    ((512, 18, 10), (9, 9, 7), 2),    # 256-point rows
    ((96, 30, 20), (9, 9, 7), 2),     # 48-point rows (runtime-plan kernel)
    ((16, 2048, 6), (5, 9, 5), 2),    # 2048-point dim-2 columns (1024-thread blocks, 128 KB LDS)
    ((20, 1024, 8), (5, 9, 5), 2),    # 1024-point dim-2 columns (512-thread blocks)
    ((24, 12, 1024), (5, 5, 9), 2),   # 1024-point dim-3 convolution columns
])
def test_omw_vs_oracle(oracle, shape, psf_shape, numit):
    import petakit5d_amd as pk
    print(f"[omw] shape={shape} psf={psf_shape} numit={numit}", flush=True)
    I, psf = _case(shape, psf_shape, shear=20.0)
    rng = np.random.default_rng(7)
    psf_b = (_flip(psf) * (1 + 0.3 * rng.standard_normal(psf.shape))).astype(np.float32)
    J, err = pk.decon_lucy_omw_function(I, psf, psf_b, numit, Background=100, save16bit=False)
    R32, _ = oracle.decon_lucy_omw_function(I, psf, psf_b, numit, Background=100, save16bit=False)
    R64, _ = oracle.decon_lucy_omw_function(I, psf, psf_b, numit, Background=100, save16bit=False, precision="f64")
    e32, e64 = rel_l2(J, R32), rel_l2(J, R64)
    print(f"  rel-L2 vs CPU FP32 {e32:.3e}  vs FP64 {e64:.3e}")
    assert e32 <= TOL


def test_omw_wiener_butterworth_backprojector_vs_oracle(oracle):
    """OMW-2 with the Wiener-Butterworth back projector (BackProjector, alpha
    0.005, order 10): it lifts high frequencies, so FP32 rounding matters more
    than with the flipped PSF; the bar stays 1e-4 against CPU FP32."""
    import petakit5d_amd as pk
    I, psf = _case((64, 64, 48), (21, 21, 31), shear=31.5)
    pb = pk.BackProjector(psf, "wiener-butterworth", 0.005, 1, 10)[0].astype(np.float32)
    J, _ = pk.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    R32, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    R64, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False, precision="f64")
    e32, e64 = rel_l2(J, R32), rel_l2(J, R64)
    print(f"[omw wb] rel-L2 vs CPU FP32 {e32:.3e}  vs FP64 {e64:.3e}  (CPU FP32 vs FP64 {rel_l2(R32, R64):.3e})")
    assert e32 <= TOL


def test_biggs15_128x128x64(oracle):
    """128x128x64, 200 emitters, PSF 31x31x41, Biggs RL 15 iterations."""
    import petakit5d_amd as pk
    from petakit5d_amd import synth
    I, psf = synth.emitter_case()
    J, _, _ = pk.decon_lucy_function(I, psf, 15, Background=100, save16bit=False)
    R32, _, _ = oracle.decon_lucy_function(I, psf, 15, Background=100, save16bit=False)
    R64, _, _ = oracle.decon_lucy_function(I, psf, 15, Background=100, save16bit=False, precision="f64")
    e32, e64 = rel_l2(J, R32), rel_l2(J, R64)
    print(f"[biggs15 128x128x64] rel-L2 vs CPU FP32 {e32:.3e}  vs FP64 {e64:.3e}")
    assert e32 <= TOL


@pytest.mark.parametrize("kw", [
    dict(save16bit=True),
    dict(save16bit=True, bbox=[3, 5, 2, 40, 33, 20]),
    dict(save16bit=False, dampFactor=1.5),
    dict(save16bit=False, scaleFactor=2.5),
    dict(save16bit=True, scaleFactor=3.0, deconOffset=7.0),
    dict(save16bit=False, EdgeErosion=2),
    dict(save16bit=False, EdgeErosion=4),   # strel('disk', 3): the periodic-line decomposition (a 5x5 square)
    dict(save16bit=True, EdgeErosion=8),    # strel('disk', 7), XR_RLdeconFrame3D.m:41's default
], ids=lambda k: "-".join(f"{a}={b}" for a, b in k.items()))
def test_epilogue_options_vs_oracle(oracle, kw):
    import petakit5d_amd as pk
    I, psf = _case((48, 40, 24), (9, 9, 11), seed=11)
    I[:3, :, :] = 0  # an irregular I ~= 0 mask for the offset / erosion branches
    J, _, _ = pk.decon_lucy_function(I, psf, 4, Background=100, **kw)
    R, _, _ = oracle.decon_lucy_function(I, psf, 4, Background=100, **kw)
    assert J.shape == R.shape and J.dtype == R.dtype
    if J.dtype == np.uint16:
        d = np.abs(J.astype(np.int64) - R.astype(np.int64))
        # uint16 rounding of values that agree to ~1e-6 relative may differ by one count
        assert d.max() <= 1 and (d > 0).mean() < 1e-3
    else:
        assert rel_l2(J, R) <= TOL


@pytest.mark.parametrize("method", ["rl", "omw"])
@pytest.mark.parametrize("bg", [99.37, 0.0])
def test_single_input_fractional_background_vs_oracle(oracle, method, bg):
    """A single-class input takes max(I - Background, 0) in single precision
    (decon_lucy_function.m:93-94, decon_lucy_omw_function.m:86-87), so a
    non-integer Background is subtracted as is, not rounded to uint16."""
    import petakit5d_amd as pk
    print(f"[single input] {method} Background={bg}", flush=True)
    I16, psf = _case((48, 40, 30), (9, 9, 11), seed=17)
    rng = np.random.default_rng(3)
    I = (I16.astype(np.float32) + rng.random(I16.shape, dtype=np.float32)).astype(np.float32, order="F")
    if method == "rl":
        J, _, _ = pk.decon_lucy_function(I, psf, 5, Background=bg, save16bit=False)
        R, _, _ = oracle.decon_lucy_function(I, psf, 5, Background=bg, save16bit=False)
    else:
        J, _ = pk.decon_lucy_omw_function(I, psf, _flip(psf), 2, Background=bg, save16bit=False)
        R, _ = oracle.decon_lucy_omw_function(I, psf, _flip(psf), 2, Background=bg, save16bit=False)
    e = rel_l2(J, R)
    print(f"  rel-L2 vs CPU FP32 {e:.3e}")
    assert e <= TOL


def test_device_resident_matches_host(oracle):
    import torch
    import petakit5d_amd as pk
    I, psf = _case((64, 48, 30), (9, 9, 11), seed=5)
    Jh, _ = pk.decon_lucy_omw_function(I, psf, _flip(psf), 2, Background=100)
    It = torch.from_numpy(np.ascontiguousarray(I.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    Jd, _ = pk.decon_lucy_omw_function(It, psf, _flip(psf), 2, Background=100)
    assert isinstance(Jd, torch.Tensor) and Jd.is_cuda
    np.testing.assert_array_equal(Jd.cpu().numpy(), Jh)


def test_psf2otf_vs_oracle(oracle):
    import petakit5d_amd as pk
    rng = np.random.default_rng(1)
    psf = rng.random((7, 6, 5)).astype(np.float32)
    got = pk.decon_psf2otf(psf, (30, 24, 25))
    ref = oracle.decon_psf2otf(psf, (30, 24, 25), precision="f64")
    assert rel_l2(got, ref) < 2e-6


def test_errors():
    import petakit5d_amd as pk
    I = np.zeros((32, 32, 16), np.uint16)
    with pytest.raises(pk.PkDeconError) as e:
        pk.decon_lucy_function(I, np.ones((33, 3, 3), np.float32), 2)
    assert e.value.status == "PK_ERR_OUT_SIZE"
    with pytest.raises(pk.PkDeconError) as e:
        pk.decon_lucy_function(np.zeros((31, 32, 16), np.uint16), np.ones((3, 3, 3), np.float32), 2)
    assert e.value.status == "PK_ERR_UNSUPPORTED_SIZE"
    with pytest.raises(TypeError):
        pk.decon_lucy_function(np.zeros((32, 32, 16)), np.ones((3, 3, 3), np.float32), 2)


@pytest.mark.parametrize("method", ["omw", "rl"])
def test_unfused_plain_path_matches_fused(method):
    """The MATLAB-structured path (one kernel per gpuArray statement) runs the
    same transforms and per-element arithmetic; the two differ only where the
    compiler contracts a product into an FMA differently in separate kernels
    (last-bit differences, rel-L2 <= 1e-6) and, for Biggs, through the order
    of the lambda reductions."""
    import petakit5d_amd as pk
    I, psf = _case((64, 48, 32), (9, 9, 11), seed=21)
    pb = _flip(psf)
    if method == "omw":
        A, _ = pk.decon_lucy_omw_function(I, psf, pb, 3, Background=100, save16bit=False)
        B, _ = pk.decon_lucy_omw_function(I, psf, pb, 3, Background=100, save16bit=False, flags=PLAIN)
        assert rel_l2(A, B) < 1e-6
    else:
        A, _, _ = pk.decon_lucy_function(I, psf, 6, Background=100, save16bit=False)
        B, _, _ = pk.decon_lucy_function(I, psf, 6, Background=100, save16bit=False, flags=PLAIN)
        assert rel_l2(A, B) < 1e-5


PLAIN = 3  # PK_FLAG_UNFUSED | PK_FLAG_NO_OTF_CACHE


def test_autotune_keeps_parity_and_batch_identity(oracle, tmp_path, monkeypatch):
    """pk_autotune measures launch shapes / pass order for a size; results stay
    within the bar, batched and single calls stay bitwise identical, and the
    choice is appended to PK_TUNE_FILE."""
    import petakit5d_amd as pk
    tf = tmp_path / "tune.txt"
    monkeypatch.setenv("PK_TUNE_FILE", str(tf))
    ctx = pk.Context(0)
    shape = (96, 80, 60)
    rep = ctx.autotune(shape)
    print(f"[autotune] {rep}")
    assert rep.startswith("96x80x60:") and "conv dim" in rep
    line = tf.read_text().split()
    assert line[:3] == ["96", "80", "60"] and line[-1] in ("2", "3")
    I, psf = _case(shape, (11, 11, 13), shear=20.0)
    pb = _flip(psf)
    J, _ = pk.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False, ctx=ctx)
    R32, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    assert rel_l2(J, R32) <= TOL
    B = pk.decon_batch([I, I[:, :, :48].copy()], psf, pb, 2, method="omw", Background=100, save16bit=False, ctx=ctx)
    np.testing.assert_array_equal(B[0], J)
    Jr, _, _ = pk.decon_lucy_function(I, psf, 4, Background=100, save16bit=False, ctx=ctx)
    Rr, _, _ = oracle.decon_lucy_function(I, psf, 4, Background=100, save16bit=False)
    assert rel_l2(Jr, Rr) <= TOL
    ctx.close()


@pytest.mark.parametrize("method", ["omw", "rl"])
def test_graph_replays_match_eager(method):
    """The iteration loop of a repeated call is captured (2nd call) and replayed
    (later calls) as a HIP graph; every replay, with other eager work and
    synchronisations in between, equals the first (eager) call bitwise."""
    import petakit5d_amd as pk
    ctx = pk.Context(0)
    I, psf = _case((64, 48, 40), (9, 9, 11), seed=31)
    I2, psf2 = _case((48, 40, 30), (7, 7, 9), seed=32)
    pb = _flip(psf)

    def run():
        if method == "omw":
            return pk.decon_lucy_omw_function(I, psf, pb, 3, Background=100, save16bit=False, ctx=ctx)[0]
        return pk.decon_lucy_function(I, psf, 5, Background=100, save16bit=False, ctx=ctx)[0]

    ref = run()  # eager
    for r in range(4):  # capture, then replays
        other = pk.decon_lucy_function(I2, psf2, 2, Background=100, save16bit=False, ctx=ctx)[0]  # eager work between
        assert np.isfinite(other).all()
        got = run()
        np.testing.assert_array_equal(got, ref, err_msg=f"replay {r}")
    ctx.close()


@pytest.mark.parametrize("n1", [72, 96, 120, 144, 160, 192, 240, 288, 480, 640, 1920,
                                90, 100, 150, 180, 216, 250, 270, 300, 324, 360, 432, 576, 1000,
                                15, 45, 75, 125, 135, 225, 405, 729])
def test_row_lengths_vs_oracle(oracle, n1):
    """Every family of compile-time row kernels (row length n1 = 2L, L = P * E)
    against the oracle, OMW-2 and Biggs-4, through the single and the batched
    entry points: every stage on home registers with P a power of two (96,
    160, 192, 640, 1920) or rows packed per wavefront with idle lanes (72: P = 6,
    100: 5, 150: 5, 216: 9, 250: 25, 1000: 50); one stage fed from LDS because
    its radix does not divide E (90: E = 9, 120: 5, 144: 8, 180: 10, 240: 8,
    270: 9, 288: 9, 300: 10, 324: 9, 360: 9, 432: 8, 480: 8, 576: 9).  Odd rows
    (full-length transform of n1 real values): 15 (P = 1), 45, 75, 125, 135,
    225, 405 with compile-time kernels, 729 on the runtime-plan kernel."""
    import petakit5d_amd as pk
    shape = (n1, 12, 10)
    print(f"[rows] n1={n1}", flush=True)
    I, psf = _case(shape, (7, 7, 5), seed=40 + n1 % 17)
    pb = _flip(psf)
    J, _ = pk.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    R, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    assert rel_l2(J, R) <= TOL
    B = pk.decon_batch([I], psf, pb, 2, method="omw", Background=100, save16bit=False)
    np.testing.assert_array_equal(B[0], J)
    Jr, _, _ = pk.decon_lucy_function(I, psf, 4, Background=100, save16bit=False)
    Rr, _, _ = oracle.decon_lucy_function(I, psf, 4, Background=100, save16bit=False)
    assert rel_l2(Jr, Rr) <= TOL


# (shape, launch shapes {xE, yE, yNT, yT, zE, zNT, zT, conv axis}): strided-axis
# plans whose values per thread miss a prime of the length, so one or two stages
# (radix 2-5) run from LDS with round-robin butterflies
MIXED_COL = [
    ((64, 60, 45), (8, 6, 256, 16, 9, 256, 16, 3)),    # y: 6, 2, then 5 from LDS; z: 9, then 5 from LDS
    ((64, 60, 45), (8, 4, 256, 16, 5, 256, 16, 2)),    # two LDS stages per axis: y 4 | 5, 3; z 5 | 3, 3
    ((64, 60, 45), (8, 12, 256, 8, 15, 256, 16, 3)),   # y: 12, then 5 from LDS (P = 5); z all-register
    ((48, 360, 40), (6, 24, 256, 16, 8, 256, 16, 3)),  # y: 24, 3, then 5 from LDS; z: 8, then 5 from LDS
    ((48, 360, 40), (6, 12, 256, 8, 5, 256, 16, 2)),   # y: 12, 6, then 5 from LDS; z: 5, then 4, 2 from LDS
]


@pytest.mark.parametrize("shape,shapes", MIXED_COL, ids=lambda v: "x".join(map(str, v)))
def test_mixed_column_plans_vs_oracle(oracle, shape, shapes):
    """Strided-axis transforms with stages fed from LDS (their radix does not
    divide the values per thread), set through pk_tune_set: OMW-2 and Biggs-4
    against the oracle, and the batched entry point bitwise equal to the
    single call under the same shapes."""
    import petakit5d_amd as pk
    print(f"[mixed cols] {shape} {shapes}", flush=True)
    ctx = pk.Context(0)
    ctx.tune_set(shape, shapes)
    I, psf = _case(shape, (9, 9, 11), seed=7)
    pb = _flip(psf)
    J, _ = pk.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False, ctx=ctx)
    R, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    assert rel_l2(J, R) <= TOL
    B = pk.decon_batch([I], psf, pb, 2, method="omw", Background=100, save16bit=False, ctx=ctx)
    np.testing.assert_array_equal(B[0], J)
    Jr, _, _ = pk.decon_lucy_function(I, psf, 4, Background=100, save16bit=False, ctx=ctx)
    Rr, _, _ = oracle.decon_lucy_function(I, psf, 4, Background=100, save16bit=False)
    assert rel_l2(Jr, Rr) <= TOL
    ctx.close()


def test_tune_set_rejects_infeasible_shapes():
    import petakit5d_amd as pk
    ctx = pk.Context(0)
    with pytest.raises(pk.PkDeconError):
        ctx.tune_set((64, 60, 45), (8, 7, 256, 16, 9, 256, 16, 3))  # 7 does not divide 60
    with pytest.raises(pk.PkDeconError):
        ctx.tune_set((64, 60, 45), (8, 6, 256, 32, 9, 256, 16, 3))  # 10 threads x 32 columns > 256
    with pytest.raises(pk.PkDeconError):
        ctx.tune_set((48, 360, 40), (6, 24, 512, 32, 8, 256, 16, 3))  # LDS-fed stages: 256-thread kernels only
    ctx.close()


EPILOGUE_PASS = 8  # PK_FLAG_EPILOGUE_PASS


@pytest.mark.parametrize("shape", [(48, 40, 24), (45, 36, 20)], ids=["even-rows", "odd-rows"])
@pytest.mark.parametrize("method,numit", [("omw", 1), ("omw", 3), ("rl", 4)])
@pytest.mark.parametrize("kw", [
    dict(save16bit=True),
    dict(save16bit=False),
    # dim-1 crop (0-based start, size): (1, 39), (3, 30) and (2, 31) store
    # single values, (0, 36) stores pairs
    dict(save16bit=True, bbox=[2, 5, 2, 40, 33, 20]),
    dict(save16bit=False, bbox=[4, 2, 3, 33, 30, 19]),
    dict(save16bit=True, bbox=[3, 2, 3, 33, 30, 19]),
    dict(save16bit=True, bbox=[1, 2, 3, 36, 30, 19]),
    dict(save16bit=False, dampFactor=1.5),
    dict(save16bit=True, scaleFactor=3.0, deconOffset=7.0),
    dict(save16bit=False, EdgeErosion=2, deconOffset=1.0),
    dict(save16bit=True, EdgeErosion=8),
], ids=lambda k: "-".join(f"{a}={b}" for a, b in k.items()).replace(" ", ""))
def test_fused_epilogue_matches_epilogue_pass(shape, method, numit, kw):
    """The last update writes the output through the epilogue itself; the
    epilogue as a pass of its own after that update must give the same bits
    for every option (crop on even and odd boundaries, damp, scale / offset /
    mask, erosion, both output classes, packed and odd-length rows)."""
    import petakit5d_amd as pk
    print(f"[fused epilogue] shape={shape} method={method} numit={numit} {kw}", flush=True)
    I, psf = _case(shape, (9, 9, 11), seed=13)
    I[:3, :, :] = 0  # an irregular I ~= 0 mask
    if method == "omw":
        A, _ = pk.decon_lucy_omw_function(I, psf, _flip(psf), numit, Background=100, **kw)
        B, _ = pk.decon_lucy_omw_function(I, psf, _flip(psf), numit, Background=100, flags=EPILOGUE_PASS, **kw)
    else:
        A, _, _ = pk.decon_lucy_function(I, psf, numit, Background=100, **kw)
        B, _, _ = pk.decon_lucy_function(I, psf, numit, Background=100, flags=EPILOGUE_PASS, **kw)
    assert A.shape == B.shape and A.dtype == B.dtype
    np.testing.assert_array_equal(A, B)


@pytest.mark.parametrize("method,numit,step", [("rl", 6, 1), ("omw", 7, 2), ("rl", 7, 2)])
def test_debug_err_mat_and_iteration_files(oracle, tmp_path, method, numit, step):
    """debug=true (decon_lucy_function.m:144-172, decon_lucy_omw_function.m:
    114-143): err_mat(k, 2) is the mean squared difference between the k-th
    estimate and the background-subtracted input -- checked against the
    oracle's estimates after k iterations --, column 3 its ratio to row 1,
    column 4 the reference's stopping criterion (worked by hand below for
    saveStep 2), trimmed as the reference trims; every saveStep-th estimate is
    written as Iter_%04d.tif (uint16)."""
    import os
    import petakit5d_amd as pk
    I, psf = _case((48, 40, 24), (9, 9, 11), seed=23)
    pb = _flip(psf)
    folder = str(tmp_path / "dbg")
    kw = dict(Background=100, save16bit=False, debug=True, debug_folder=folder, saveStep=step)
    if method == "rl":
        J, err, _ = pk.decon_lucy_function(I, psf, numit, **kw)
    else:
        J, err = pk.decon_lucy_omw_function(I, psf, pb, numit, **kw)
    Ibg = np.maximum(I.astype(np.float32) - np.float32(100), 0).astype(np.float64)
    mse = []
    for k in range(1, numit + 1):
        if method == "rl":
            Rk, _, _ = oracle.decon_lucy_function(I, psf, k, Background=100, save16bit=False)
        else:
            Rk, _ = oracle.decon_lucy_omw_function(I, psf, pb, k, Background=100, save16bit=False)
        mse.append(float(np.mean((Rk.astype(np.float64) - Ibg) ** 2)))
        if k % step == 0:
            f = os.path.join(folder, f"Iter_{k:04d}.tif")
            T = pk.readtiff(f)
            assert T.dtype == np.uint16
            assert np.abs(T.astype(np.int64) - np.clip(np.floor(Rk.astype(np.float64) + 0.5), 0, 65535)).max() <= 1
    rows = numit // step if method == "omw" else max(k for k in range(1, numit + 1)
                                                       if k % step == 0 and k > 2 * step) // step
    assert err.shape == (rows, 4)
    np.testing.assert_array_equal(err[:, 0], np.arange(1, rows + 1))
    np.testing.assert_allclose(err[:, 1], mse[:rows], rtol=1e-4)
    np.testing.assert_allclose(err[:, 2], err[:, 1] / err[0, 1], rtol=1e-12)
    if step == 1:  # rows 1, 2: 1; row k > 2: min(|e3(k) - e3(k-1)| / 10, |e3(k) + e3(k-2) - 2 e3(k-1)| / 2)
        e3 = err[:, 2]
        want = [1.0, 1.0] + [min(abs(e3[k] - e3[k - 1]) / 10, abs(e3[k] + e3[k - 2] - 2 * e3[k - 1]) / 2)
                             for k in range(2, rows)]
        np.testing.assert_allclose(err[:, 3], want, rtol=1e-12)
    else:  # saveStep 2, NUMIT 7: rows 1..3 come from iterations 1..3, all before 3 * saveStep: criterion 1
        np.testing.assert_array_equal(err[:, 3], np.ones(rows))


def test_debug_rl_needs_three_save_steps(tmp_path):
    """decon_lucy_function.m:170-172 reads istp, which only an iteration k > 2
    saveStep that is a multiple of saveStep defines."""
    import petakit5d_amd as pk
    I, psf = _case((32, 32, 16), (7, 7, 9), seed=24)
    with pytest.raises(RuntimeError, match="istp"):
        pk.decon_lucy_function(I, psf, 2, Background=100, debug=True, debug_folder=str(tmp_path / "d"))


@pytest.mark.parametrize("shape,psf_shape,kind", [
    ((15, 6, 8), (1, 1, 1), "delta"),     # a one-voxel PSF on a small volume of odd 15-point rows
    ((72, 12, 10), (5, 5, 3), "zeros"),   # an all-zero image: I ./ max(CX, eps) keeps J at 0
    ((45, 6, 1), (3, 3, 1), "ramp"),      # a single plane of odd rows
])
def test_small_and_degenerate_inputs_vs_oracle(oracle, shape, psf_shape, kind):
    """Edge cases of the decon functions against the oracle, RL (Biggs, 5
    iterations) and OMW (2 iterations), at row and column lengths whose
    kernels the other tests run: a one-voxel PSF (OTF = 1, so J stays
    max(I - bg, 0)), an all-zero image, and a single plane."""
    import petakit5d_amd as pk
    from petakit5d_amd import synth
    print(f"[edge] {shape} psf {psf_shape} {kind}", flush=True)
    n = int(np.prod(shape))
    if kind == "zeros":
        I = np.zeros(shape, np.uint16)
    else:
        I = np.asfortranarray((np.arange(n).reshape(shape, order="F") * 37 % 500 + 120).astype(np.uint16))
    psf = np.ones(psf_shape, np.float32) if kind == "delta" else synth.gaussian_psf(psf_shape, 0.8, 0.8)
    pb = _flip(psf)
    for method, it in (("rl", 5), ("omw", 2)):
        if method == "rl":
            J, _, _ = pk.decon_lucy_function(I, psf, it, Background=100, save16bit=False)
            R, _, _ = oracle.decon_lucy_function(I, psf, it, Background=100, save16bit=False)
        else:
            J, _ = pk.decon_lucy_omw_function(I, psf, pb, it, Background=100, save16bit=False)
            R, _ = oracle.decon_lucy_omw_function(I, psf, pb, it, Background=100, save16bit=False)
        assert J.shape == shape and np.isfinite(J).all()
        if kind == "zeros":
            assert not J.any() and not R.any()
        else:
            assert rel_l2(J, R) <= TOL, (method, rel_l2(J, R))
        if kind == "delta":  # OTF = 1: every iteration keeps J = max(I - bg, 0)
            np.testing.assert_allclose(J, np.maximum(I.astype(np.float32) - 100, 0), rtol=1e-6, atol=1e-4)

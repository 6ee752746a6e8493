"""matlab/pk_decon_mex.c compiled against a stand-in mx API
(tests/mex_harness) and linked to libpkdecon.so: dispatch, argument checks
and error identifiers run without MATLAB; a decon call reaches the C-ABI
(PK_ERR_NO_DEVICE on a host without a GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mex_gateway(tmp_path):
    exe = tmp_path / "mexdrv"
    lib = os.path.join(ROOT, "petakit5d_amd")
    cmd = ["gcc", "-O1", "-Wall", "-I", os.path.join(ROOT, "tests", "mex_harness"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "matlab", "pk_decon_mex.c"), os.path.join(ROOT, "tests", "mex_harness", "mx_fake.c"),
           "-L", lib, "-lpkdecon", f"-Wl,-rpath,{lib}", "-o", str(exe)]
    subprocess.check_call(cmd)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout.splitlines()
    assert out[0] == "fgfn 108" and out[1] == "fgfn 270"
    assert out[2] == "error pkdecon:invalidArg"
    assert out[3] == "error pkdecon:unsupportedType"
    # fourth call reaches the engine: no device here, a shape-checked result on a GPU host
    assert out[4] in ("error pkdecon:noDevice", "rl ok 8 8 4")
    # BackProjector runs on the host: the same answer with or without a device
    assert out[5] == "wbbp ok 9 9 9 1"
    assert out[6] == "error pkdecon:invalidArg"


def _build(tmp_path):
    exe = tmp_path / "mexdrv"
    lib = os.path.join(ROOT, "petakit5d_amd")
    subprocess.check_call(["gcc", "-O1", "-Wall", "-I", os.path.join(ROOT, "tests", "mex_harness"), "-I",
                           os.path.join(ROOT, "include"), os.path.join(ROOT, "matlab", "pk_decon_mex.c"),
                           os.path.join(ROOT, "tests", "mex_harness", "mx_fake.c"), "-L", lib, "-lpkdecon",
                           f"-Wl,-rpath,{lib}", "-o", str(exe)])
    return exe


@pytest.mark.gpu
def test_mex_gateway_on_device_vs_oracle(tmp_path, oracle):
    """The MEX gateway on the device: 'rl' (Biggs, 6 iterations) and 'omw'
    (2 iterations) on a uint16 volume with an options struct (Background 100,
    save16bit false), the MEX-allocated outputs against the oracle's CPU
    single-precision path at rel-L2 <= 1e-4; err_mat is NUMIT x 4 zeros
    (debug off, decon_lucy_function.m:117)."""
    import numpy as np
    from conftest import rel_l2
    from petakit5d_amd import synth
    exe = _build(tmp_path)
    shape, pshape = (64, 48, 30), (11, 11, 13)
    psf = synth.gaussian_psf(pshape, 1.2, 3.0, 20.0)
    img = synth._fft_convolve_circular(synth.point_emitters(shape, 50, (500, 3000), synth.rng(81, 0)), psf) + 100
    I = synth.poisson_u16(img, synth.rng(81, 1))
    pb = np.asfortranarray(psf[::-1, ::-1, ::-1])
    for name, a in (("I.u16", I), ("psf.f32", psf), ("psfb.f32", pb)):
        np.asfortranarray(a).ravel(order="F").tofile(str(tmp_path / name))
    r = subprocess.run([str(exe), "files", str(tmp_path), *map(str, shape), *map(str, pshape), "6", "2", "100"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.stdout.splitlines() == ["rl ok 64 48 30", "omw ok 64 48 30", "reset ok"]
    J_rl = np.fromfile(str(tmp_path / "J_rl.f32"), np.float32).reshape(shape, order="F")
    J_omw = np.fromfile(str(tmp_path / "J_omw.f32"), np.float32).reshape(shape, order="F")
    err = np.fromfile(str(tmp_path / "err_rl.f64"), np.float64)
    R_rl, _, _ = oracle.decon_lucy_function(I, psf, 6, Background=100, save16bit=False)
    R_omw, _ = oracle.decon_lucy_omw_function(I, psf, pb, 2, Background=100, save16bit=False)
    e_rl, e_omw = rel_l2(J_rl, R_rl), rel_l2(J_omw, R_omw)
    print(f"[mex] rl rel-L2 {e_rl:.3e}  omw rel-L2 {e_omw:.3e}")
    assert e_rl <= 1e-4 and e_omw <= 1e-4
    assert err.shape == (24,) and not err.any()

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

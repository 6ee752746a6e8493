"""The library's host code under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/build.sh: every source compiled for
the host only, with an empty offload bundle in place of device code).

pk_host_san drives the TIFF LZW and Zarr LZ4 codecs (round trips, every short
output capacity, truncated, bit-flipped and random streams), BackProjector's
five types and three resolution rules, the size rules (findGoodFactorNumber,
deskew, deskew-rotate on extreme geometry, strel('disk', r) for r <= 255) and
the no-device / NULL-context error paths.  mexdrv_san runs the MEX gateway
(matlab/pk_decon_mex.c) on the stand-in mx API, whose registry frees what a
MEX call leaves behind as MATLAB does, so any leak reported is the gateway's
or the library's.  Runs only where no GPU is visible: the build registers no
kernels and expects pk_ctx_create to find no device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "tests", "sanitize", "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def sanitized():
    if shutil.which("hipcc") is None:
        pytest.skip("hipcc not found")
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("host-only sanitizer build: runs where no GPU is visible")
    subprocess.run(["bash", os.path.join(ROOT, "tests", "sanitize", "build.sh")], check=True, timeout=900,
                   capture_output=True)
    return OUT


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout.splitlines()


def test_host_code_under_sanitizers(sanitized):
    out = _run(os.path.join(sanitized, "pk_host_san"))
    assert out == ["ok lzw", "ok lz4", "ok backprojector", "ok sizes", "ok contexts", "all ok"]


def test_mex_gateway_under_sanitizers(sanitized):
    out = _run(os.path.join(sanitized, "mexdrv_san"))
    assert out == ["fgfn 108", "fgfn 270", "error pkdecon:invalidArg", "error pkdecon:unsupportedType",
                   "error pkdecon:noDevice", "wbbp ok 9 9 9 1", "error pkdecon:invalidArg"]

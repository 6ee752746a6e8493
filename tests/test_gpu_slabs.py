"""Slab-by-slab row passes (PK_SLAB_PLANES: the dim-1 and dim-2 passes run a
few z planes at a time between the whole-volume dim-3 convolutions) give
bitwise the same output as whole-volume passes: every row is transformed with
the same arithmetic whichever slab holds it.  uint16 input, so the slabs also
walk the uint16 copy of the estimate.  The setting is read once per process,
so the slab run is a child process (one at a time, started, not exec'd)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
from test_gpu_slabs import run_cases
out = run_cases()
np.savez(sys.argv[2], **out)
print("slab child done", flush=True)
"""


def _case(shape, seed):
    from petakit5d_amd import synth
    psf = synth.gaussian_psf((11, 11, 13), 1.2, 3.0, 15.0)
    g0, g1 = synth.rng(seed, 0), synth.rng(seed, 1)
    obj = synth.point_emitters(shape, 40, (500, 3000), g0)
    img = synth._fft_convolve_circular(obj, psf) + 100.0
    return synth.poisson_u16(img, g1), psf


def run_cases():
    """OMW-2 and Biggs-3 on two uint16 volumes (conv along dim 3, the built-in
    order); float outputs, and the ratio passes' launch count."""
    import petakit5d_amd as pk
    ctx = pk.Context(0)
    ctx.profile(True)
    out = {}
    for i, shape in enumerate([(64, 64, 32), (96, 40, 48)]):
        I, psf = _case(shape, 21 + i)
        psf_b = np.ascontiguousarray(psf[::-1, ::-1, ::-1])
        out[f"omw{i}"], _ = pk.decon_lucy_omw_function(I, psf, psf_b, 2, Background=100, save16bit=False, ctx=ctx)
        out[f"rl{i}"], _, _ = pk.decon_lucy_function(I, psf, 3, Background=100, save16bit=False, ctx=ctx)
    out["ratio_launches"] = np.array(ctx.profile_read()["x_ratio"][0])
    ctx.close()
    return out


def test_slab_passes_match_whole_volume(tmp_path):
    print("[slabs] whole volume in this process", flush=True)
    ref = run_cases()
    path = tmp_path / "slabs.npz"
    env = dict(os.environ, PK_SLAB_PLANES="8")
    print("[slabs] PK_SLAB_PLANES=8 in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, str(path)], env=env, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = np.load(path)
    print(f"  x_ratio launches: whole volume {int(ref['ratio_launches'])}, slabs {int(got['ratio_launches'])}")
    assert int(got["ratio_launches"]) > int(ref.pop("ratio_launches"))  # the slab path ran
    for k, v in ref.items():
        assert v.dtype == np.float32 and np.isfinite(v).all() and v.max() > 0
        assert np.array_equal(v.view(np.uint32), got[k].view(np.uint32)), k

"""omw_backprojector_generation with the reference's signature
(microscopeDataProcessing/decon/omw_backprojector_generation.m:1-166), served
by libpkdecon.so (host toolbox steps + GPU nearest-edge search)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PkBpOpts, i64x3

_NAMES = {"otfcumthresh": "OTFCumThresh", "otfareathresh": "OTFAreaThresh", "hannwinbounds": "hannWinBounds",
          "minintthrsh": "minIntThrsh", "ctx": "ctx"}


def omw_backprojector_generation(psf, alpha, skewed=None, **kwargs):
    """[b_omw, OTF_bp_omw, abs_OTF_c, OTF_mask] = omw_backprojector_generation(psf, alpha, skewed, ...)

    skewed=None (MATLAB []) auto-detects skewed-space data.  Name-value
    options match case-insensitively: OTFCumThresh (0.9), OTFAreaThresh (100),
    hannWinBounds ([0.8 1]), minIntThrsh (1e-3)."""
    o = {"OTFCumThresh": 0.9, "OTFAreaThresh": 100, "hannWinBounds": (0.8, 1.0), "minIntThrsh": 1e-3, "ctx": None}
    for k, v in kwargs.items():
        name = _NAMES.get(k.lower())
        if name is None:
            raise TypeError(f"'{k}' is not a recognized parameter")
        o[name] = v
    hw = np.asarray(o["hannWinBounds"], dtype=np.float64).reshape(-1)
    if hw.size != 2:
        raise ValueError("hannWinBounds must have two elements")
    p = np.asarray(psf)
    if p.ndim > 3:
        raise ValueError("PSF must be at most 3-D")
    d = tuple(int(x) for x in p.shape) + (1,) * (3 - p.ndim)
    pf = np.asfortranarray(p.astype(np.float32).reshape(d, order="F"))
    opts = PkBpOpts(float(o["OTFCumThresh"]), float(o["OTFAreaThresh"]), float(hw[0]), float(hw[1]),
                    float(o["minIntThrsh"]))
    b = np.zeros(d, np.float32, order="F")
    omw = np.zeros(d, np.complex64, order="F")
    absc = np.zeros(d, np.float32, order="F")
    mask = np.zeros(d, np.float32, order="F")
    used = ctypes.c_int32(0)
    sk = -1 if skewed is None or (np.size(skewed) == 0) else (1 if bool(skewed) else 0)
    ctx = o["ctx"] or _lib.default_context()
    ctx.check(_lib.load().pk_omw_backprojector(
        ctx.handle, ctypes.c_void_p(pf.ctypes.data), i64x3(d), float(alpha), sk, ctypes.byref(opts),
        ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(omw.ctypes.data), ctypes.c_void_p(absc.ctypes.data),
        ctypes.c_void_p(mask.ctypes.data), ctypes.byref(used)))
    m = mask > 0 if used.value == 0 else mask.astype(np.float64)
    return b, omw, absc, m

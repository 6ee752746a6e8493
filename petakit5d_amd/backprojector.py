"""Back projectors with the reference's signatures, served by libpkdecon.so:
omw_backprojector_generation (microscopeDataProcessing/decon/
omw_backprojector_generation.m:1-166; host toolbox steps + GPU nearest-edge
search) and BackProjector (third_parties/WBDeconvolution/BackProjector.m:1-233;
host)."""
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


BP_TYPES = {"traditional": 0, "gaussian": 1, "butterworth": 2, "wiener": 3, "wiener-butterworth": 4}


def BackProjector(PSF_fp, bp_type="traditional", alpha=0.001, beta=1, n=10, resFlag=1, iRes=(0, 0, 0),
                  verboseFlag=0, ctx=None, info=False):
    """[PSF_bp, OTF_bp] = BackProjector(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes, verboseFlag)

    Defaults are the reference's by argument count (BackProjector.m:28-63):
    alpha 0.001, beta 1, n 10, resFlag 1, iRes [0 0 0].  alpha == 1 / beta == 1
    take the traditional back projector's cutoff gain.  PSF_bp is double and
    OTF_bp complex double, as in the reference; verboseFlag only printed and
    plotted there and is accepted and ignored.  info=True also returns the
    dict {FWHM, cutoff, beta_fp, alpha, beta, beta_wienerx}."""
    t = BP_TYPES.get(str(bp_type).lower())
    if t is None:
        raise ValueError("bp_type does not match any back-projector type")
    p = np.asarray(PSF_fp)
    if p.ndim > 3:
        raise ValueError("PSF must be at most 3-D")
    d = tuple(int(x) for x in p.shape) + (1,) * (3 - p.ndim)
    pf = np.asfortranarray(p.astype(np.float32).reshape(d, order="F"))
    out = np.zeros(d, np.float64, order="F")
    otf = np.zeros(d, np.complex128, order="F")
    nfo = (ctypes.c_double * 10)()
    ires = (ctypes.c_double * 3)(*[float(x) for x in np.asarray(iRes, np.float64).reshape(-1)[:3]])
    h = ctx.handle if ctx is not None else None
    st = _lib.load().pk_wb_backprojector(h, ctypes.c_void_p(pf.ctypes.data), i64x3(d), t, float(alpha), float(beta),
                                         int(n), int(resFlag), ires, ctypes.c_void_p(out.ctypes.data),
                                         ctypes.c_void_p(otf.ctypes.data), nfo)
    if st != 0:
        msg = _lib.load().pk_last_error(h).decode(errors="replace") if h else "BackProjector failed"
        raise _lib.PkDeconError(st, msg)
    if info:
        v = list(nfo)
        return out, otf, {"FWHM": v[0:3], "cutoff": v[3:6], "beta_fp": v[6], "alpha": v[7], "beta": v[8],
                          "beta_wienerx": v[9]}
    return out, otf

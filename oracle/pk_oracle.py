"""ctypes front end of the C oracle (oracle/pk_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for petakit5d_amd's HIP path.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.

Arrays use MATLAB semantics: numpy arrays whose shape is the MATLAB size
(n1, n2, n3), stored column-major (Fortran order), dim 1 contiguous.

Functions mirror the reference (paths relative to the PetaKit5D tree):
  fftn / ifftn            MATLAB builtins used at decon_lucy_function.m:136,142
  decon_psf2otf           microscopeDataProcessing/decon/decon_psf2otf.m:1-46
  decon_lucy_function     microscopeDataProcessing/decon/decon_lucy_function.m:1-215 (useGPU=false)
  decon_lucy_omw_function microscopeDataProcessing/decon/decon_lucy_omw_function.m:1-178 (useGPU=false)
  decon_mask_edge_erosion microscopeDataProcessing/decon/decon_mask_edge_erosion.m:1-30
  strel_disk              strel('disk', r) as decon_mask_edge_erosion.m:22-27 uses it
  findGoodFactorNumber    utils/findGoodFactorNumber.m:1-26
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpkoracle.so")
_lib = None


class _Opts(ctypes.Structure):
    _fields_ = [
        ("background", ctypes.c_double),
        ("dampFactor", ctypes.c_double),
        ("scaleFactor", ctypes.c_double),
        ("deconOffset", ctypes.c_double),
        ("edgeErosion", ctypes.c_int32),
        ("save16bit", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        p = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        for suf in ("f32", "f64"):
            getattr(L, f"pko_fftn_{suf}").argtypes = [p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
            getattr(L, f"pko_psf2otf_{suf}").argtypes = [p, i64p, i64p, p]
            getattr(L, f"pko_decon_lucy_{suf}").argtypes = [p, i64p, p, i64p, ctypes.c_int, ctypes.POINTER(_Opts), p, p]
            getattr(L, f"pko_decon_omw_{suf}").argtypes = [p, i64p, p, i64p, p, i64p, ctypes.c_int,
                                                           ctypes.POINTER(_Opts), p, p]
        L.pko_mask_edge_erosion.argtypes = [p, i64p, ctypes.c_int]
        L.pko_strel_disk.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pko_strel_disk.restype = ctypes.c_int
        L.pko_find_good_factor_number.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.pko_find_good_factor_number.restype = ctypes.c_int64
        L.pko_to_u16.argtypes = [ctypes.c_double]
        L.pko_to_u16.restype = ctypes.c_uint16
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


_ERRS = {1: "invalid argument", 2: "outSizeIsSmallerThanPsfSize", 3: "out of memory"}


def _check(rc):
    if rc != 0:
        raise OracleError(_ERRS.get(rc, f"oracle error {rc}"))


def _dims3(shape):
    s = tuple(int(v) for v in shape)
    if len(s) > 3:
        raise ValueError("at most 3-D")
    return s + (1,) * (3 - len(s))


def _i64(v):
    return (ctypes.c_int64 * 3)(*v)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def fftn(x, inverse=False, precision="f32"):
    """MATLAB fftn / ifftn (ifftn scaled by 1/N) of a 1-3-D array."""
    x = np.asarray(x)
    shp = x.shape
    d = _dims3(shp)
    ct = np.complex128 if precision == "f64" else np.complex64
    buf = np.asfortranarray(x.astype(ct).reshape(d, order="F"))
    _check(getattr(lib(), f"pko_fftn_{precision}")(_ptr(buf), d[0], d[1], d[2], 1 if inverse else 0))
    return buf.reshape(shp, order="F")


def ifftn(x, precision="f32"):
    return fftn(x, inverse=True, precision=precision)


def decon_psf2otf(psf, outSize=None, precision="f32"):
    """decon_psf2otf(psf, outSize) -> full complex OTF of size outSize."""
    rt = np.float64 if precision == "f64" else np.float32
    psf = np.asfortranarray(np.asarray(psf, dtype=rt))
    pd = _dims3(psf.shape)
    od = pd if outSize is None else _dims3(outSize)
    ct = np.complex128 if precision == "f64" else np.complex64
    out = np.zeros(od, dtype=ct, order="F")
    _check(getattr(lib(), f"pko_psf2otf_{precision}")(_ptr(psf), _i64(pd), _i64(od), _ptr(out)))
    return out


def _opts(Background, dampFactor, scaleFactor, deconOffset, EdgeErosion, save16bit):
    return _Opts(float(0 if Background is None else Background), float(dampFactor), float(scaleFactor),
                 float(deconOffset), int(EdgeErosion), 1 if save16bit else 0)


def _prep_input(I, precision):
    rt = np.float64 if (precision == "f64" or np.asarray(I).dtype == np.float64) else np.float32
    return np.asfortranarray(np.asarray(I).astype(rt)), rt


def _crop(J, bbox):
    if bbox is None or len(bbox) == 0:
        return J
    b = [int(v) for v in bbox]
    return J[b[0] - 1:b[3], b[1] - 1:b[4], b[2] - 1:b[5]]


def decon_lucy_function(I, PSF, NUMIT, Background=None, save16bit=True, dampFactor=1, scaleFactor=1,
                        deconOffset=0, EdgeErosion=0, bbox=None, precision="f32"):
    """CPU branch of decon_lucy_function (single(I), or double for f64).

    Returns (J, err_mat, k) like the reference (err_mat zeros, debug off)."""
    Ic, rt = _prep_input(I, precision)
    suf = _suffix(rt)
    psf = np.asfortranarray(np.asarray(PSF, dtype=rt))
    d, pd = _dims3(Ic.shape), _dims3(psf.shape)
    J = np.zeros(d, dtype=rt, order="F")
    u16 = np.zeros(d, dtype=np.uint16, order="F") if save16bit else None
    o = _opts(Background, dampFactor, scaleFactor, deconOffset, EdgeErosion, save16bit)
    _check(getattr(lib(), f"pko_decon_lucy_{suf}")(_ptr(Ic), _i64(d), _ptr(psf), _i64(pd), int(NUMIT),
                                                   ctypes.byref(o), _ptr(J),
                                                   _ptr(u16) if u16 is not None else None))
    out = u16 if save16bit else J
    return _crop(out.reshape(Ic.shape, order="F"), bbox), np.zeros((int(NUMIT), 4)), int(NUMIT)


def decon_lucy_omw_function(I, PSF_f, PSF_b, NUMIT, Background=None, save16bit=True, dampFactor=1,
                            scaleFactor=1, deconOffset=0, EdgeErosion=0, bbox=None, precision="f32"):
    """CPU branch of decon_lucy_omw_function.  Returns (J, err_mat)."""
    Ic, rt = _prep_input(I, precision)
    suf = _suffix(rt)
    pf = np.asfortranarray(np.asarray(PSF_f, dtype=rt))
    pb = np.asfortranarray(np.asarray(PSF_b, dtype=rt))
    d = _dims3(Ic.shape)
    J = np.zeros(d, dtype=rt, order="F")
    u16 = np.zeros(d, dtype=np.uint16, order="F") if save16bit else None
    o = _opts(Background, dampFactor, scaleFactor, deconOffset, EdgeErosion, save16bit)
    _check(getattr(lib(), f"pko_decon_omw_{suf}")(_ptr(Ic), _i64(d), _ptr(pf), _i64(_dims3(pf.shape)), _ptr(pb),
                                                  _i64(_dims3(pb.shape)), int(NUMIT), ctypes.byref(o), _ptr(J),
                                                  _ptr(u16) if u16 is not None else None))
    out = u16 if save16bit else J
    return _crop(out.reshape(Ic.shape, order="F"), bbox), np.zeros((int(NUMIT), 4))


def decon_mask_edge_erosion(mask, EdgeErosion):
    m = np.asfortranarray((np.asarray(mask) > 0).astype(np.uint8))
    d = _dims3(m.shape)
    _check(lib().pko_mask_edge_erosion(_ptr(m), _i64(d), int(EdgeErosion)))
    return m.astype(bool)


def strel_disk(r):
    """getnhood(strel('disk', r)) (default N = 4) as a boolean (2R+1) x (2R+1)
    array, rows = dim 1."""
    w = (ctypes.c_int * (2 * int(r) + 1))()
    R = lib().pko_strel_disk(int(r), w)
    if R < 0:
        raise OracleError("invalid argument")
    nh = np.zeros((2 * R + 1, 2 * R + 1), bool)
    for i in range(2 * R + 1):
        nh[i, R - w[i]:R + w[i] + 1] = True
    return nh


def findGoodFactorNumber(givenNum, directionFlag=1, allowOddSeven=False, maxPrime=7):
    return int(lib().pko_find_good_factor_number(int(givenNum), int(directionFlag), 1 if allowOddSeven else 0,
                                                 int(maxPrime)))


def to_uint16(x):
    """MATLAB uint16(x) for an array (round half away from zero, saturate, NaN->0)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        f = np.floor(x)
        r = f + ((x - f) >= 0.5)  # exact for x >= 0; negatives clip to 0 below
        r = np.where(np.isnan(x), 0.0, np.clip(r, 0, 65535))
    return r.astype(np.uint16)

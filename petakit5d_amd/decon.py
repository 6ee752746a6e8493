"""MATLAB-signature front end of the MI355X deconvolution engine.

Mirrors the reference's call surface (paths relative to the PetaKit5D tree):

  decon_lucy_function(I, PSF, NUMIT, Name=Value...)       decon/decon_lucy_function.m:1-215
  decon_lucy_omw_function(I, PSF_f, PSF_b, NUMIT, ...)    decon/decon_lucy_omw_function.m:1-178
  decon_psf2otf(psf, outSize)                             decon/decon_psf2otf.m:1-46
  findGoodFactorNumber(n, directionFlag, allowOddSeven)   utils/findGoodFactorNumber.m:1-26
  crop_psf_to_data(psf, imSize)                           decon/RLdecon.m:264-270, 326-332

Arrays carry MATLAB semantics: the array shape is the MATLAB size
(n1, n2, n3) and memory is column-major (dim 1 contiguous).  numpy inputs
(host) give numpy outputs; torch CUDA tensors (device-resident) give torch
CUDA tensors -- a tensor is accepted when its strides are column-major,
e.g. ``t.permute(2, 1, 0)`` of a contiguous (n3, n2, n1) tensor.

Name-value options are matched case-insensitively like MATLAB's
inputParser (CaseSensitive=false).  Engine extensions beside the reference's
options: pad (zero-pad non-2/3/5-smooth sizes), ctx (device context), flags
(PK_FLAG_* comparison baselines).  All arithmetic runs in libpkdecon.so on
the GPU; ``useGPU=False`` is rejected because this engine has no CPU branch.
"""
from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import _lib
from ._lib import PK_MEM_DEVICE, PK_MEM_HOST, PK_SINGLE, PK_UINT16, PkDeconError, PkOpts, PkVol, i64x3

_OPTION_NAMES = ["Background", "useGPU", "save16bit", "dampFactor", "scaleFactor", "deconOffset", "EdgeErosion",
                 "bbox", "debug", "debug_folder", "saveStep", "pad", "ctx", "flags"]
_DEFAULTS = {"Background": None, "useGPU": True, "save16bit": True, "dampFactor": 1, "scaleFactor": 1,
             "deconOffset": 0, "EdgeErosion": 0, "bbox": None, "debug": False, "debug_folder": "./debug/",
             "saveStep": 1, "pad": False, "ctx": None, "flags": 0}


def _parse(kwargs: dict) -> dict:
    lut = {n.lower(): n for n in _OPTION_NAMES}
    out = dict(_DEFAULTS)
    for k, v in kwargs.items():
        name = lut.get(k.lower())
        if name is None:
            raise TypeError(f"'{k}' is not a recognized parameter")
        out[name] = v
    return out


def _is_torch(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor)


def _dims3(shape) -> tuple:
    s = tuple(int(v) for v in shape)
    if len(s) > 3:
        raise ValueError("volumes are at most 3-D")
    return s + (1,) * (3 - len(s))


def _fortran_strides(d):
    return (1, d[0], d[0] * d[1])


class _Vol:
    """A pk_vol plus the Python object that owns its memory."""

    def __init__(self, obj, dtype_code, mem, dims):
        self.obj = obj
        self.dims = dims
        ptr = obj.data_ptr() if mem == PK_MEM_DEVICE else obj.ctypes.data
        self.c = PkVol(ctypes.c_void_p(ptr), dtype_code, mem, (ctypes.c_int64 * 3)(*dims))


def _input_vol(I):
    if _is_torch(I):
        import torch
        if not I.is_cuda:
            I = I.numpy()
        else:
            d = _dims3(I.shape)
            t = I.reshape(d) if I.dim() < 3 else I
            if t.dtype == torch.uint16:
                code = PK_UINT16
            elif t.dtype == torch.float32:
                code = PK_SINGLE
            else:
                raise TypeError(f"input class {t.dtype} not supported (uint16 or single)")
            if tuple(t.stride()) != _fortran_strides(d):
                t = t.permute(2, 1, 0).contiguous().permute(2, 1, 0)
            return _Vol(t, code, PK_MEM_DEVICE, d), "torch"
    a = np.asarray(I)
    if a.dtype == np.uint16:
        code = PK_UINT16
    elif a.dtype == np.float32:
        code = PK_SINGLE
    else:
        raise TypeError(f"input class {a.dtype} not supported by the GPU engine (uint16 or single)")
    d = _dims3(a.shape)
    a = np.asfortranarray(a.reshape(d, order="F"))
    return _Vol(a, code, PK_MEM_HOST, d), "numpy"


def _output_vol(kind, dims, save16bit, device):
    if kind == "torch":
        import torch
        dt = torch.uint16 if save16bit else torch.float32
        t = torch.empty((dims[2], dims[1], dims[0]), dtype=dt, device=device).permute(2, 1, 0)
        return _Vol(t, PK_UINT16 if save16bit else PK_SINGLE, PK_MEM_DEVICE, dims)
    a = np.empty(dims, dtype=np.uint16 if save16bit else np.float32, order="F")
    return _Vol(a, PK_UINT16 if save16bit else PK_SINGLE, PK_MEM_HOST, dims)


def _psf_array(P):
    if _is_torch(P):
        P = P.detach().cpu().numpy()
    a = np.asarray(P)
    if a.size == 0:
        raise ValueError("empty PSF")
    d = _dims3(a.shape)
    return np.asfortranarray(a.astype(np.float32, copy=False).reshape(d, order="F")), d


def _smooth235(n: int) -> bool:
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def _pad_to_smooth(vol: _Vol, kind: str):
    """Zero post-pad every axis to the next 2/3/5-smooth length."""
    d = vol.dims
    pd = tuple(findGoodFactorNumber(n, 1, False, maxPrime=5) for n in d)
    if pd == d:
        return vol
    if kind == "torch":
        import torch
        src = vol.obj
        t = torch.zeros((pd[2], pd[1], pd[0]), dtype=src.dtype, device=src.device).permute(2, 1, 0)
        t[: d[0], : d[1], : d[2]] = src
        return _Vol(t, vol.c.dtype, PK_MEM_DEVICE, pd)
    a = np.zeros(pd, dtype=vol.obj.dtype, order="F")
    a[: d[0], : d[1], : d[2]] = vol.obj
    return _Vol(a, vol.c.dtype, PK_MEM_HOST, pd)


def _opts_struct(o: dict, orig_dims, padded: bool) -> tuple:
    s = PkOpts()
    _lib.load().pk_opts_default(ctypes.byref(s))
    bg = o["Background"]
    s.background = 0.0 if bg is None or np.size(bg) == 0 else float(np.asarray(bg).reshape(-1)[0])
    s.damp_factor = float(o["dampFactor"])
    s.scale_factor = float(o["scaleFactor"])
    s.decon_offset = float(o["deconOffset"])
    s.edge_erosion = int(o["EdgeErosion"])
    s.save16bit = 1 if o["save16bit"] else 0
    s.flags = int(o.get("flags", 0))  # PK_FLAG_* comparison baselines (engine extension)
    bbox = o["bbox"]
    if bbox is not None and np.size(bbox) > 0:
        b = [int(v) for v in np.asarray(bbox).reshape(-1)]
        if len(b) != 6:
            raise ValueError("bbox must be [s1 s2 s3 e1 e2 e3]")
    elif padded:
        b = [1, 1, 1, orig_dims[0], orig_dims[1], orig_dims[2]]
    else:
        b = None
    if b is not None:
        s.has_bbox = 1
        for i in range(6):
            s.bbox[i] = b[i]
        od = (b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1)
    else:
        od = tuple(orig_dims)
    return s, od


def _run(method: int, I, psfs, NUMIT: int, kwargs: dict):
    o = _parse(kwargs)
    if not o["useGPU"]:
        raise ValueError("useGPU=false: this engine implements the GPU branch only")
    ctx = o["ctx"] or _lib.default_context()
    vol, kind = _input_vol(I)
    orig = vol.dims
    padded = False
    if not all(_smooth235(n) for n in orig):
        if not o["pad"]:
            raise PkDeconError(3, f"size {orig} is not 2/3/5-smooth; pass pad=True to zero-pad to "
                                  "findGoodFactorNumber sizes (the border then differs from an unpadded FFT)")
        vol = _pad_to_smooth(vol, kind)
        padded = True
    opts, od = _opts_struct(o, orig, padded)
    device = vol.obj.device if kind == "torch" else None
    out = _output_vol(kind, od, bool(o["save16bit"]), device)
    if kind == "torch":
        import torch
        torch.cuda.current_stream(device).synchronize()
    L = _lib.load()
    err = np.zeros((max(int(NUMIT), 0), 4), dtype=np.float64, order="F")
    errp = ctypes.c_void_p(err.ctypes.data) if err.size else None
    arrs = [_psf_array(p) for p in psfs]
    hook = None
    if o["debug"]:
        opts.flags |= PK_FLAG_DEBUG
        hook = _debug_hook(str(o["debug_folder"]), orig, int(NUMIT))
        ctx.check(L.pk_set_iter_hook(ctx.handle, int(o["saveStep"]), hook, None))
    if method == 0:
        (pf, fd), = arrs
        st = L.pk_decon_lucy(ctx.handle, ctypes.byref(vol.c), ctypes.c_void_p(pf.ctypes.data), i64x3(fd),
                             int(NUMIT), ctypes.byref(opts), ctypes.byref(out.c), errp)
    else:
        (pf, fd), (pb, bd) = arrs
        st = L.pk_decon_lucy_omw(ctx.handle, ctypes.byref(vol.c), ctypes.c_void_p(pf.ctypes.data), i64x3(fd),
                                 ctypes.c_void_p(pb.ctypes.data), i64x3(bd), int(NUMIT), ctypes.byref(opts),
                                 ctypes.byref(out.c), errp)
    if hook is not None:
        L.pk_set_iter_hook(ctx.handle, 1, None, None)
    ctx.check(st)
    if o["debug"]:
        err = _debug_trim(err, method, int(NUMIT), int(o["saveStep"]))
    J = out.obj
    if len(np.shape(I)) < 3 and kind == "numpy":
        J = J.reshape(J.shape[: len(np.shape(I))], order="F")
    return J, err


PK_FLAG_DEBUG = 16
_HOOK_T = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(ctypes.c_int64))


def _debug_hook(folder, orig, numit):
    """The iteration hook of debug=true: writes uint16(J) of every saveStep-th
    iteration as folder/Iter_%04d.tif (decon_lucy_function.m:159-167), cropped
    to the caller's size when the engine padded the volume."""
    import os

    from .tiff_io import writetiff
    if numit >= 1:
        os.makedirs(folder, exist_ok=True)  # mkdir(debug_folder) at k == 1

    def cb(_user, k, data, dims):
        d = (int(dims[0]), int(dims[1]), int(dims[2]))
        a = np.ctypeslib.as_array(data, shape=(d[0] * d[1] * d[2],)).reshape(d, order="F")
        a = a[:orig[0], :orig[1], :orig[2]].astype(np.float64)
        r = np.sign(a) * np.floor(np.abs(a) + 0.5)  # uint16(): round half away from zero, saturate
        writetiff(np.asfortranarray(np.clip(np.nan_to_num(r), 0, 65535).astype(np.uint16)),
                  os.path.join(folder, f"Iter_{int(k):04d}.tif"))
    return _HOOK_T(cb)


def _debug_trim(err, method, numit, estep):
    """The reference's last debug step: decon_lucy_function keeps rows
    1..istp, istp = k / estep of the last k > 2 estep that is a multiple of
    estep (MATLAB raises an undefined-variable error when there is none,
    decon_lucy_function.m:170-172); decon_lucy_omw_function keeps rows
    1..floor(NUMIT / estep) (decon_lucy_omw_function.m:141-143)."""
    if method == 1:
        return np.asfortranarray(err[:numit // estep])
    ks = [k for k in range(1, numit + 1) if k % estep == 0 and k > 2 * estep]
    if not ks:
        raise RuntimeError("Unrecognized function or variable 'istp' (decon_lucy_function.m:171: debug needs "
                           "NUMIT >= 3 * saveStep)")
    return np.asfortranarray(err[:ks[-1] // estep])


def decon_lucy_function(I, PSF, NUMIT, **kwargs: Any):
    """[J, err_mat, k] = decon_lucy_function(I, PSF, NUMIT, Name=Value...)

    Biggs-accelerated Richardson-Lucy (decon_lucy_function.m:1-215)."""
    J, err = _run(0, I, [PSF], int(NUMIT), kwargs)
    return J, err, int(NUMIT)


def decon_lucy_omw_function(I, PSF_f, PSF_b, NUMIT, **kwargs: Any):
    """[J, err_mat] = decon_lucy_omw_function(I, PSF_f, PSF_b, NUMIT, Name=Value...)

    RL with the OTF-masked Wiener back projector (decon_lucy_omw_function.m:1-178)."""
    return _run(1, I, [PSF_f, PSF_b], int(NUMIT), kwargs)


def decon_psf2otf(psf, outSize=None, ctx=None):
    """otf = decon_psf2otf(psf, outSize): full complex64 OTF (decon_psf2otf.m:1-46)."""
    p, pd = _psf_array(psf)
    od = pd if outSize is None else _dims3(outSize)
    out = np.zeros(od, dtype=np.complex64, order="F")
    ctx = ctx or _lib.default_context()
    ctx.check(_lib.load().pk_psf2otf(ctx.handle, ctypes.c_void_p(p.ctypes.data), i64x3(pd), i64x3(od),
                                     ctypes.c_void_p(out.ctypes.data)))
    return out


def findGoodFactorNumber(givenNum, directionFlag=1, allowOddSeven=False, maxPrime=7):
    """findGoodFactorNumber.m:1-26; maxPrime=5 gives the 2/3/5-smooth variant."""
    return _lib.find_good_factor_number(int(givenNum), int(directionFlag), bool(allowOddSeven), int(maxPrime))


def crop_psf_to_data(psf, imSize):
    """Centred crop of a PSF larger than the data (RLdecon.m:264-270):
    s = max(0, floor((pSize - imSize)/2)) + 1, t = s + min(pSize, imSize) - 1."""
    p = np.asarray(psf)
    ps = _dims3(p.shape)
    im = _dims3(imSize)
    if not any(a > b for a, b in zip(ps, im)):
        return p
    s = [max(0, (a - b) // 2) for a, b in zip(ps, im)]
    t = [s[i] + min(ps[i], im[i]) for i in range(3)]
    p3 = p.reshape(ps, order="F")
    return p3[s[0]:t[0], s[1]:t[1], s[2]:t[2]]


def fft_r2c(x):
    """Half-spectrum R2C of a column-major float32 CUDA tensor (diagnostic)."""
    import torch
    d = _dims3(x.shape)
    if tuple(x.stride()) != _fortran_strides(d):
        x = x.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    h1 = d[0] // 2 + 1
    out = torch.empty((d[2], d[1], h1), dtype=torch.complex64, device=x.device).permute(2, 1, 0)
    ctx = _lib.default_context()
    ctx.check(_lib.load().pk_fft_r2c(ctx.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                     i64x3(d)))
    return out


def fft_c2r(spec, dims):
    """Unnormalised C2R of a column-major complex64 CUDA half spectrum (diagnostic)."""
    import torch
    d = _dims3(dims)
    h1 = d[0] // 2 + 1
    if tuple(spec.shape) != (h1, d[1], d[2]):
        raise ValueError("spectrum shape mismatch")
    if tuple(spec.stride()) != _fortran_strides((h1, d[1], d[2])):
        spec = spec.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    out = torch.empty((d[2], d[1], d[0]), dtype=torch.float32, device=spec.device).permute(2, 1, 0)
    ctx = _lib.default_context()
    ctx.check(_lib.load().pk_fft_c2r(ctx.handle, ctypes.c_void_p(spec.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                     i64x3(d)))
    return out

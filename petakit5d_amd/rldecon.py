"""Per-volume decon driver with the semantics of RLdecon.m for in-memory data
(rawdata given, PSF given as an array): PSF crop to the data, OMW back
projector generation with an on-disk cache, method dispatch, background,
scale, offset, damp, bbox and edge erosion passed to the decon functions.

Files: the back projector is cached as a TIFF under the reference's file name
(RLdecon.m:292), with the OTF magnitude and OTF mask beside it (:313-323);
rawdata may be a TIFF or Zarr path, read as readtiff / readzarr would
(RLdecon.m:160-185).  Deviation: the engine pads non-2/3/5-smooth sizes with
zeros and crops back (pad=True)."""
from __future__ import annotations

import hashlib
import math
import os

import numpy as np

from .backprojector import omw_backprojector_generation
from .psfgen import psf_gen_new
from .tiff_io import readtiff, writetiff
from .decon import _is_torch, crop_psf_to_data, decon_lucy_function, decon_lucy_omw_function

__all__ = ["RLdecon_volume", "omw_bp_filename", "psfgen_crop", "psfgen_dz_ratio"]

_BP_CACHE: dict = {}


def omw_bp_filename(psfgenPath: str, psfFsn: str, wienerAlpha: float, OTFCumThresh: float) -> str:
    """Back projector cache name:
    sprintf('%s/%s_back_projector_alpha_%0.6f_otf_cum_thresh_%0.6f.tif', ...) (RLdecon.m:292)."""
    return os.path.join(psfgenPath, f"{psfFsn}_back_projector_alpha_{wienerAlpha:0.6f}_otf_cum_thresh_"
                                    f"{OTFCumThresh:0.6f}.tif")


def _omw_bp(psf, wienerAlpha, skewed, OTFCumThresh, hannWinBounds, psfgenPath, psfFsn, ctx):
    """OMW back projector of psf (RLdecon.m:290-325): loaded from the cache file
    when it exists, else generated (omw_backprojector_generation) and written
    with the OTF magnitude and mask files; also kept in memory per (PSF bytes,
    alpha, threshold, window, skewed)."""
    key = (hashlib.sha1(np.ascontiguousarray(psf).tobytes()).hexdigest(), psf.shape, float(wienerAlpha),
           float(OTFCumThresh), tuple(float(x) for x in hannWinBounds), None if skewed is None else bool(skewed))
    if key in _BP_CACHE:
        return _BP_CACHE[key]
    fn = omw_bp_filename(psfgenPath, psfFsn, wienerAlpha, OTFCumThresh) if psfgenPath and psfFsn else None
    if fn and os.path.exists(fn):
        b = readtiff(fn).astype(np.float32, copy=False)
    else:
        b, _, absc, mask = omw_backprojector_generation(psf, wienerAlpha, skewed, OTFCumThresh=OTFCumThresh,
                                                        hannWinBounds=list(hannWinBounds), ctx=ctx)
        b = np.asarray(b, dtype=np.float32)
        if fn:
            os.makedirs(psfgenPath, exist_ok=True)
            tag = f"{os.getpid()}"
            # each file written aside, then moved in place (RLdecon.m:310-323)
            for final, data in ((fn, b),
                                (os.path.join(psfgenPath, f"{psfFsn}_OTF.tif"), np.asarray(absc, np.float32)),
                                (os.path.join(psfgenPath, f"{psfFsn}_OTF_mask_otf_cum_thresh_{OTFCumThresh:0.6f}.tif"),
                                 np.asarray(mask).astype(np.uint8))):
                tmp = final[:-4] + f"_{tag}.tif"
                writetiff(data, tmp)
                os.replace(tmp, final)
    _BP_CACHE[key] = b
    return b


def psfgen_dz_ratio(SkewAngle, objectiveScan=False):
    """dz_ratio = sind(abs(SkewAngle)) for stage scans, 1 for objective scans
    (RLdecon.m:101-106); sind is exact at multiples of 90 degrees."""
    if objectiveScan:
        return 1.0
    a = abs(float(SkewAngle)) % 360.0
    exact = {0.0: 0.0, 90.0: 1.0, 180.0: 0.0, 270.0: -1.0}
    return exact[a] if a in exact else math.sin(math.radians(a))


def _to_uint16(a):
    """MATLAB uint16(single): round half away from zero, saturate to [0, 65535],
    NaN -> 0.  A CUDA tensor stays on its device (torch uint16 result)."""
    if _is_torch(a):
        import torch
        t = a.to(torch.float32)
        r = torch.sign(t) * torch.floor(t.abs() + 0.5)
        return torch.nan_to_num(r, nan=0.0).clamp_(0, 65535).to(torch.int32).to(torch.uint16)
    a = np.asarray(a, dtype=np.float32)
    r = np.sign(a) * np.floor(np.abs(a) + np.float32(0.5))
    return np.asfortranarray(np.clip(np.nan_to_num(r), 0, 65535).astype(np.uint16))


def _read_volume(path):
    ext = os.path.splitext(path.rstrip("/"))[1].lower()
    if ext in (".tif", ".tiff"):
        return readtiff(path)
    if ext == ".zarr":
        from .zarr_io import read_zarr
        return read_zarr(path)
    raise ValueError(f"input file '{path}': .tif, .tiff or .zarr expected")


def psfgen_crop(psf):
    """Crop of a generated PSF to its non-zero extent plus a one-voxel margin,
    the same on both sides of every axis so the centre stays put
    (RLdecon.m:250-260)."""
    p = np.asarray(psf)
    ps = np.array(p.shape[:3])
    crop = []
    for ax in range(3):
        other = tuple(i for i in range(3) if i != ax)
        nzi = np.nonzero(p.sum(axis=other))[0]
        if nzi.size == 0:
            crop.append(0)
            continue
        crop.append(min(int(nzi[0]), int(ps[ax] - 1 - nzi[-1])) - 1)
    c = np.maximum(0, np.array(crop))
    return p[c[0]:ps[0] - c[0], c[1]:ps[1] - c[1], c[2]:ps[2] - c[2]]


def RLdecon_volume(rawdata, psf, nIter, RLMethod="simplified", Background=0.0, wienerAlpha=0.005,
                   OTFCumThresh=0.9, hannWinBounds=(0.8, 1.0), skewed=None, flipZstack=False, save16bit=True,
                   dampFactor=1.0, scaleFactor=None, deconOffset=0.0, deconBbox=None, EdgeErosion=0,
                   psfgenPath=None, psfFsn=None, psfGen=False, dzPSF=None, dz=None, Deskew=False, SkewAngle=32.45,
                   xyPixelSize=0.108, Reverse=False, Crop=False, objectiveScan=False, Rotate=False,
                   DSRCombined=True, xStepThresh=2.0, Resample=None, nTapering=0, interpMethod="linear",
                   fixIter=True, errThresh=None, ErodeMaskfile=None, debug=False, saveStep=5, debug_folder=None,
                   useGPU=True, ctx=None):
    """deconvolved = RLdecon(..., 'rawdata', rawdata, ...) for one volume
    (rawdata an array, or a .tif / .zarr path; psf an array or a .tif path).

    Follows RLdecon.m: PSF as single (:235), centre-cropped when larger than the
    data (:264-270); for 'omw' the back projector from the cache or
    omw_backprojector_generation (:290-325), cropped to the data (:326-332);
    flipZstack (:346-348); with nIter > 0 and any non-zero voxel, scaleFactor
    [] -> 1 and PSF ./ sum(PSF) (:367-373), then decon_lucy_function
    ('simplified') or decon_lucy_omw_function ('omw') with Background,
    save16bit, dampFactor, scaleFactor, deconOffset, bbox = deconBbox and
    EdgeErosion (:387-404); otherwise the raw data, cropped to deconBbox
    (:415-420).  psfGen: the PSF first goes through psf_gen_new(psf, dzPSF,
    dz * dz_ratio, 1.5, 'masked'), dz_ratio = sind(|SkewAngle|) for stage
    scans and 1 for objectiveScan (:101-106, 247), and, for methods other than
    'omw', psfgen_crop (:250-260); with psfgenPath and psfFsn the generated
    PSF is loaded from psfgenPath/psfFsn_RLMethod.tif when that exists
    (:219-229) and written there after generation for file inputs
    (:272-279).  Deskew / Rotate return (deconvolved, ds, dsr) (None for a
    step not run): ds = deskewFrame3D(single(deconvolved), SkewAngle, dz,
    xyPixelSize, Reverse, 'Crop', Crop) (:448-449), dsr = rotateFrame3D(ds,
    SkewAngle, zAniso, Reverse, 'Crop', true) with zAniso = dz_ratio dz /
    xyPixelSize (:107, 483-486), or of the deconvolved volume itself for
    objective scans (:494-497); uint16 when save16bit (:464-468, 498-501).
    Deskew and Rotate of a stage scan with DSRCombined (the default) run
    deskewRotateFrame3D(deconvolved, SkewAngle, dz, xyPixelSize, 'reverse',
    Reverse, 'Crop', true, 'resample', Resample, 'xStepThresh', xStepThresh,
    'save16bit', save16bit) instead and return ds = None (:488-493).
    objectiveScan turns Deskew off (:147-149); Rotate of a stage scan without
    Deskew produces nothing in the reference and is an error here.  Other
    methods ('original' deconvlucy, 'cudagen') are not this engine's.

    ErodeMaskfile (:364-413): with EdgeErosion > 0 and an existing Zarr mask
    the decon runs without erosion and its result is multiplied by mask > 0.
    debug / saveStep (:374-380) go to the decon functions, with saveStep =
    min(saveStep, nIter) and the iteration files in debug_folder (the
    reference's deconPath/<name>_debug/; a temporary folder when not given).
    fixIter and errThresh are parsed and unused, as in the reference
    (:84-85).  interpMethod: the reference validates it with @isnumeric
    (:26), so only the default is usable and a method name is rejected.
    nTapering > 0 (edgetaper of each slice, :352-357) is not served."""
    if RLMethod not in ("simplified", "omw"):
        raise ValueError(f"RLMethod '{RLMethod}': this engine serves 'simplified' and 'omw'")
    if str(interpMethod) != "linear":
        raise ValueError("interpMethod: RLdecon.m:26 validates it with @isnumeric, so a method name is rejected; "
                         "call deskewFrame3D / rotateFrame3D / deskewRotateFrame3D for 'cubic'")
    if nTapering and nTapering > 0:
        raise NotImplementedError("nTapering > 0 (edgetaper of each slice, RLdecon.m:352-357) is not served")
    if not useGPU:
        raise ValueError("useGPU=false: this engine implements the GPU branch only")
    SkewAngle = abs(float(SkewAngle))  # RLdecon.m:101
    if objectiveScan:
        Deskew = False
    if Deskew or Rotate:
        if dz is None:
            raise ValueError("Deskew / Rotate need dz")
        if Rotate and not objectiveScan and not Deskew:
            raise ValueError("Rotate of a stage scan needs Deskew (RLdecon.m:437-501 produce no rotated output "
                             "without it)")
        J = RLdecon_volume(rawdata, psf, nIter, RLMethod, Background, wienerAlpha, OTFCumThresh, hannWinBounds,
                           skewed, flipZstack, save16bit, dampFactor, scaleFactor, deconOffset, deconBbox,
                           EdgeErosion, psfgenPath, psfFsn, psfGen, dzPSF, dz, SkewAngle=SkewAngle,
                           objectiveScan=objectiveScan, ErodeMaskfile=ErodeMaskfile, debug=debug, saveStep=saveStep,
                           debug_folder=debug_folder, ctx=ctx)
        from .deskew import deskewFrame3D
        from .rotate import rotateFrame3D
        zAniso = psfgen_dz_ratio(SkewAngle, objectiveScan) * dz / xyPixelSize
        if Deskew and Rotate and DSRCombined:  # RLdecon.m:488-493
            from .deskew_rotate import deskewRotateFrame3D
            dsr = deskewRotateFrame3D(J, SkewAngle, dz, xyPixelSize, reverse=Reverse, Crop=True,
                                      objectiveScan=objectiveScan, resampleFactor=Resample, xStepThresh=xStepThresh,
                                      save16bit=save16bit, ctx=ctx)
            return J, None, (_to_uint16(dsr) if save16bit else dsr)
        ds = deskewFrame3D(J, SkewAngle, dz, xyPixelSize, Reverse, Crop=Crop, ctx=ctx) if Deskew else None
        dsr = None
        if Rotate:
            dsr = rotateFrame3D(J if objectiveScan else ds, SkewAngle, zAniso, Reverse, Crop=True,
                                ObjectiveScan=objectiveScan, ctx=ctx)
        if save16bit:
            ds = None if ds is None else _to_uint16(ds)
            dsr = None if dsr is None else _to_uint16(dsr)
        return J, ds, dsr
    file_input = isinstance(rawdata, (str, os.PathLike))
    if file_input:  # input file (RLdecon.m:335-343)
        rawdata = _read_volume(str(rawdata))
    if isinstance(psf, (str, os.PathLike)):  # PSF file (RLdecon.m:214-236)
        if psfFsn is None:
            psfFsn = os.path.splitext(os.path.basename(str(psf)))[0]
        psf = readtiff(str(psf))
    psf = np.asarray(psf.detach().cpu().numpy() if _is_torch(psf) else psf, dtype=np.float32)
    gen_fn = os.path.join(psfgenPath, f"{psfFsn}_{RLMethod}.tif") if (psfGen and psfgenPath and psfFsn) else None
    if gen_fn and os.path.exists(gen_fn):  # generated before (RLdecon.m:224-229)
        psf = readtiff(gen_fn).astype(np.float32, copy=False)
        psfGen = False
    if psfGen:
        if dzPSF is None or dz is None:
            raise ValueError("psfGen needs dzPSF and dz")
        psf = psf_gen_new(psf, dzPSF, dz * psfgen_dz_ratio(SkewAngle, objectiveScan), 1.5, "masked")
        if RLMethod != "omw":
            psf = np.ascontiguousarray(psfgen_crop(psf))
    imSize = tuple(int(n) for n in (rawdata.shape if _is_torch(rawdata) else np.shape(rawdata)))[:3]
    psf = np.asfortranarray(crop_psf_to_data(psf, imSize))
    if gen_fn and psfGen and file_input and not os.path.exists(gen_fn):  # RLdecon.m:272-279
        os.makedirs(psfgenPath, exist_ok=True)
        tmp = gen_fn[:-4] + f"_{os.getpid()}.tif"
        writetiff(psf, tmp)
        os.replace(tmp, gen_fn)
    psf_b = None
    if RLMethod == "omw":
        psf_b = np.asfortranarray(crop_psf_to_data(
            _omw_bp(psf, wienerAlpha, skewed, OTFCumThresh, hannWinBounds, psfgenPath, psfFsn, ctx), imSize))
    if flipZstack:
        rawdata = rawdata.flip(2) if _is_torch(rawdata) else np.asfortranarray(np.asarray(rawdata)[:, :, ::-1])
    has_signal = bool((rawdata != 0).any()) if _is_torch(rawdata) else bool(np.any(rawdata))
    if not (nIter > 0 and has_signal):
        out = rawdata
        if deconBbox is not None and len(deconBbox) == 6:
            b = [int(v) for v in deconBbox]
            out = out[b[0] - 1:b[3], b[1] - 1:b[4], b[2] - 1:b[5]]
        return out
    if scaleFactor is None:
        scaleFactor = 1.0
    psf = np.asfortranarray(psf / np.float32(psf.sum(dtype=np.float32)))
    use_mask = bool(EdgeErosion and EdgeErosion > 0 and ErodeMaskfile and os.path.exists(str(ErodeMaskfile)))
    kw = dict(Background=Background, save16bit=save16bit, dampFactor=dampFactor, scaleFactor=scaleFactor,
              deconOffset=deconOffset, EdgeErosion=0 if use_mask else EdgeErosion, pad=True, ctx=ctx)
    if debug:
        import tempfile
        kw.update(debug=True, saveStep=min(int(saveStep), int(nIter)),
                  debug_folder=debug_folder or tempfile.mkdtemp(prefix="RLdecon_debug_"))
    if deconBbox is not None and len(deconBbox) == 6:
        kw["bbox"] = list(deconBbox)
    if RLMethod == "simplified":
        J, _, _ = decon_lucy_function(rawdata, psf, int(nIter), **kw)
    else:
        J, _ = decon_lucy_omw_function(rawdata, psf, psf_b, int(nIter), **kw)
    if use_mask:  # deconvolved .* cast(mask > 0, class(deconvolved)) (RLdecon.m:409-413)
        from .zarr_io import read_zarr
        keep = np.asarray(read_zarr(str(ErodeMaskfile))) > 0
        if keep.shape != tuple(J.shape):
            raise ValueError(f"ErodeMaskfile holds {keep.shape}, the deconvolved volume is {tuple(J.shape)}")
        if _is_torch(J):
            import torch
            k = torch.from_numpy(np.ascontiguousarray(keep.transpose(2, 1, 0))).to(J.device).permute(2, 1, 0)
            if J.dtype == torch.uint16:  # same bits as int16; uint16 has no masked_fill
                J.view(torch.int16).masked_fill_(~k, 0)
            else:
                J = J * k.to(J.dtype)
        else:
            J = np.asfortranarray(J * keep.astype(J.dtype))
    return J

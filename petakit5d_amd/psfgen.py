"""PSF preparation before decon: background removal, masking of the PSF's own
connected component, peak centring and axial resampling, as psf_gen_new.m does
(decon/psf_gen_new.m:1-140).  Host code (numpy / scipy), run once per PSF.

Toolbox steps are restated with scipy: medfilt3 (symmetric border padding) ->
ndimage.median_filter(mode='reflect'), medfilt2 (zero padding) ->
mode='constant'; bwareaopen / bwlabeln with 26-connectivity (8 in 2-D) ->
ndimage.label with a full 3x3(x3) structure; imclose with strel('sphere', 3)
or strel('disk', 3) (the decomposed disk, a 5 x 5 square) -> binary dilation
(zero border) then erosion (border padded with true, as imerode pads binary
images)."""
from __future__ import annotations

import numpy as np
from scipy import ndimage

__all__ = ["psf_gen_new"]


def _mround(x):
    a = np.asarray(x, dtype=np.float64)
    return (np.sign(a) * np.floor(np.abs(a) + 0.5)).astype(np.int64)


def _argmax_f(a):
    """Subscripts (0-based) of MATLAB's max(a(:)): the first maximum in column-major order."""
    i = int(np.argmax(np.asarray(a).ravel(order="F")))
    return np.unravel_index(i, a.shape, order="F")


def _sphere(r):
    g = np.arange(-r, r + 1)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return (x * x + y * y + z * z) <= r * r


def _medfilt3(a):
    """medfilt3(a, [3 3 3]): 3x3x3 median, symmetric border padding (medfilt3's default)."""
    return ndimage.median_filter(a, size=3, mode="reflect")


def _medfilt2(a):
    """medfilt2(a, [3 3]): 3x3 median, zero border padding (medfilt2's default)."""
    return ndimage.median_filter(a, size=3, mode="constant", cval=0.0)


def _strel_disk(r):
    """Neighbourhood of strel('disk', r) with the default N = 4: the exact
    disk x^2 + y^2 <= r^2 for r < 3; otherwise the Minkowski sum of the four
    periodic lines along (1,0), (1,1), (0,1), (-1,1), each with 2 rp + 1
    points, rp = floor(k / |v|), k = 2r / (cot(pi/8) + 1/sin(pi/8)) (Adams'
    radial decomposition, as strel documents it), plus a horizontal and a
    vertical line of length 2 (r - e - 1) + 1 when that is at least 3, e the
    row extent of the sum (the compensation for its undersize).  r = 3 gives
    a 5 x 5 square, r = 5 the documented 9 x 9 element."""
    g = np.arange(-r, r + 1)
    if r < 3:
        y, x = np.meshgrid(g, g, indexing="ij")
        return (x * x + y * y) <= r * r
    k = 2 * r / (1 / np.tan(np.pi / 8) + 1 / np.sin(np.pi / 8))
    G = 2 * r + 1
    nh = np.zeros((G, G), bool)
    nh[r, r] = True
    for vy, vx in ((1, 0), (1, 1), (0, 1), (-1, 1)):
        rp = int(np.floor(k / np.hypot(vy, vx)))
        line = np.zeros((2 * rp * abs(vy) + 1, 2 * rp * abs(vx) + 1), bool)
        for i in range(-rp, rp + 1):
            line[rp * abs(vy) + i * vy, rp * abs(vx) + i * vx] = True
        nh = ndimage.binary_dilation(nh, structure=line)
    rows = np.nonzero(nh.any(axis=1))[0]
    e = int(rows.max() - r)
    n = 2 * (r - e - 1) + 1
    if n >= 3:
        nh = ndimage.binary_dilation(nh, structure=np.ones((1, n), bool))
        nh = ndimage.binary_dilation(nh, structure=np.ones((n, 1), bool))
    ys, xs = np.nonzero(nh)
    R = int(max(np.abs(ys - r).max(), np.abs(xs - r).max()))
    return nh[r - R:r + R + 1, r - R:r + R + 1]


def _component_mask(score, peak_map, min_size, conn_full, close_se):
    """bwareaopen(score > 0, min_size) -> bwlabeln -> L == L(peak of peak_map)
    -> imclose(., close_se)  (psf_gen_new.m:33-38, 49-55)."""
    bw = score > 0
    struct = np.ones((3,) * bw.ndim, dtype=bool) if conn_full else None
    lab, n = ndimage.label(bw, structure=struct)
    if n:
        sizes = np.bincount(lab.ravel())
        keep = sizes >= min_size
        keep[0] = False
        bw = keep[lab]
    lab, _ = ndimage.label(bw, structure=struct)
    pk = _argmax_f(peak_map)
    bw = lab == lab[pk]  # the peak's component (the background when the peak lies outside every component)
    bw = ndimage.binary_dilation(bw, structure=close_se, border_value=0)
    return ndimage.binary_erosion(bw, structure=close_se, border_value=1)


def psf_gen_new(psf, dz_psf, dz_data, medFactor=1.5, PSFGenMethod="masked"):
    """psf = psf_gen_new(psf, dz_psf, dz_data, medFactor, PSFGenMethod)
    (psf_gen_new.m:1-140): single precision in and out, MATLAB size order."""
    psf = np.asarray(psf, dtype=np.float32)
    if psf.ndim == 2:
        psf = psf[:, :, None]
    psf_raw = psf.copy()
    ny, nx, nz = psf.shape
    method = PSFGenMethod.lower()
    if ny < 10:  # psf([1:10, end-9:end], ...) indexes past the end in the reference (psf_gen_new.m:50, 59)
        raise ValueError("psf_gen_new needs at least 10 rows (dim 1)")
    # rows [1:10, end-9:end]: for ny < 20 the two ranges overlap and MATLAB keeps the repeats
    ends = np.r_[0:10, ny - 10:ny]
    if nz > 1:
        fl = psf_raw[:, :, np.r_[0:5, nz - 5:nz]]
        if np.any(fl > 0):
            if method == "median":
                psf_raw = psf_raw - np.float32(medFactor) * np.float32(np.median(fl[fl > 0]))
                psf_raw = np.maximum(psf_raw, 0)
                med = _medfilt3(psf_raw.astype(np.float64))
                bw = _component_mask(med, med, 300, True, _sphere(3))
                psf_raw = psf_raw * bw
            elif method == "masked":
                med = _medfilt3(psf)
                edge = med[ends, :, :]
                a = np.sqrt(np.abs(edge - np.float32(100))).max(axis=0, keepdims=True) * np.float32(3) + \
                    edge.mean(axis=0, keepdims=True, dtype=np.float32)
                bw = _component_mask(med - a, med, 300, True, _sphere(3))
                bg = psf_raw[ends, :, :].mean(axis=(0, 2), keepdims=True, dtype=np.float32)
                psf_raw = psf_raw * bw - bg
                psf_raw[psf_raw < 0] = 0
            else:
                raise ValueError(f"PSFGenMethod '{PSFGenMethod}'")
    else:  # 2-D PSF (psf_gen_new.m:58-86)
        p2 = psf_raw[:, :, 0]
        if method == "median":
            # psf_gen_new.m:71 filters with medfilt(psf_raw, [3, 3]), a function MATLAB does not
            # define (medfilt2 / medfilt3 exist), so the reference stops here with an error
            raise ValueError("psf_gen_new: the 2-D 'median' method calls medfilt, which is undefined in the "
                             "reference (psf_gen_new.m:71); use 'masked'")
        elif method == "masked":
            med = _medfilt2(psf[:, :, 0])
            edge = med[ends, :]
            a = np.sqrt(edge).max(axis=0, keepdims=True) * np.float32(3) + edge.mean(axis=0, keepdims=True,
                                                                                  dtype=np.float32)
            bw = _component_mask(med - a, med, 50, True, _strel_disk(3))
            bg = psf_raw[ends, :, :].mean(axis=(0, 2), dtype=np.float32)[None, :]
            p2 = p2 * bw - bg
            p2[p2 < 0] = 0
        else:
            raise ValueError(f"PSFGenMethod '{PSFGenMethod}'")
        psf_raw = p2[:, :, None].astype(np.float32)

    # centre the peak, padding z when the signal reaches too far to one side (psf_gen_new.m:88-112)
    py, px, pz = (int(v) + 1 for v in _argmax_f(psf_raw))  # 1-based, as MATLAB's ind2sub
    zsum = psf_raw.sum(axis=(0, 1))
    zst = np.nonzero(zsum)[0] + 1
    if zst.size:
        zs, zt = int(zst[0]), int(zst[-1])
        if pz - zs > nz / 2:
            add = pz - zs - int(_mround((nz - 1) / 2)) + 3
            psf_raw = np.concatenate([psf_raw, np.zeros((ny, nx, add), np.float32)], axis=2)
            nz = psf_raw.shape[2]
        if zt - pz > nz / 2:
            add = zt - pz - int(_mround((nz - 1) / 2)) + 3
            psf_raw = np.concatenate([np.zeros((ny, nx, add), np.float32), psf_raw], axis=2)
            pz = pz + add
            nz = psf_raw.shape[2]
    shift = _mround((np.array([ny, nx, nz]) + 1) / 2 - np.array([py, px, pz]))
    cropped = np.roll(psf_raw, tuple(int(s) for s in shift), axis=(0, 1, 2))

    # axial resampling by truncating the spectrum (psf_gen_new.m:114-134); new_nz / 2 is uint16
    # arithmetic (rounds to nearest), so the two halves share a plane when new_nz is odd
    ny, nx, nz = cropped.shape
    dz_ratio = dz_data / dz_psf
    if dz_ratio > 1:
        f = np.fft.fftn(cropped.astype(np.float64))
        new_nz = int(_mround(nz / dz_ratio))
        half = int(_mround(new_nz / 2))
        tr = np.zeros((ny, nx, new_nz), dtype=np.complex128)
        tr[:, :, :half] = f[:, :, :half]
        tr[:, :, new_nz - half:] = f[:, :, nz - half:]
        out = np.real(np.fft.ifftn(tr)).astype(np.float32)
    else:
        out = cropped
    return np.maximum(out, 0).astype(np.float32)

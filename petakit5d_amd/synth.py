"""Seeded synthetic microscopy volumes for the deconvolution workloads.

All arrays use MATLAB semantics: shape = MATLAB size (n1, n2, n3), Fortran
(column-major) memory.  Generation runs on the host in float64; only the
final volume is rounded to uint16 (Poisson counts) and PSFs to float32.

PSF sampling puts the peak at 0-based index floor(p/2) on every axis, the
origin decon_psf2otf.m:24 shifts to element (1,1,1), so a symmetric PSF has no
half-voxel offset.
"""
from __future__ import annotations

import numpy as np

try:
    import scipy.fft as _sfft
    import scipy.ndimage as _ndi
except ImportError:  # pragma: no cover - scipy is present in this image
    _sfft = None
    _ndi = None


def rng(config_id: int, stream: int = 0) -> np.random.Generator:
    """Independent seeded sub-stream `stream` of workload `config_id`."""
    return np.random.default_rng([0x5EED0000 + int(config_id), int(stream)])


def gaussian_psf(size, sigma_xy: float, sigma_z: float, shear_deg: float = 0.0) -> np.ndarray:
    """Anisotropic Gaussian PSF of the given size, unit sum (normalised in
    float64, then cast to float32).  shear_deg shifts slice z along dim 2 by
    tan(shear) * (z - z0) voxels (a skewed light-sheet PSF)."""
    p1, p2, p3 = (int(s) for s in size)
    c = np.array([p1 // 2, p2 // 2, p3 // 2], dtype=np.float64)
    y = np.arange(p1, dtype=np.float64)[:, None, None] - c[0]
    x = np.arange(p2, dtype=np.float64)[None, :, None] - c[1]
    z = np.arange(p3, dtype=np.float64)[None, None, :] - c[2]
    xs = x - np.tan(np.deg2rad(shear_deg)) * z
    g = np.exp(-(y ** 2 + xs ** 2) / (2 * sigma_xy ** 2) - z ** 2 / (2 * sigma_z ** 2))
    g /= g.sum()
    return np.asfortranarray(g.astype(np.float32))


def _fft_convolve_circular(vol: np.ndarray, psf: np.ndarray) -> np.ndarray:
    """Circular convolution of vol with psf centred at floor(p/2) (float64)."""
    d = vol.shape
    pad = np.zeros(d, dtype=np.float64)
    p = psf.astype(np.float64)
    pad[: p.shape[0], : p.shape[1], : p.shape[2]] = p
    pad = np.roll(pad, shift=tuple(-(s // 2) for s in p.shape), axis=(0, 1, 2))
    if _sfft is not None:
        F = _sfft.rfftn(vol, workers=-1) * _sfft.rfftn(pad, workers=-1)
        return _sfft.irfftn(F, s=d, workers=-1)
    F = np.fft.rfftn(vol) * np.fft.rfftn(pad)
    return np.fft.irfftn(F, s=d)


def point_emitters(shape, n: int, amp_range, g: np.random.Generator) -> np.ndarray:
    vol = np.zeros(shape, dtype=np.float64, order="F")
    idx = [g.integers(0, s, size=n) for s in shape]
    amps = g.uniform(amp_range[0], amp_range[1], size=n)
    np.add.at(vol, tuple(idx), amps)
    return vol


def gaussian_blobs(shape, n: int, g: np.random.Generator, sigma_range=(0.5, 2.0), lognorm=(6.0, 0.8),
                   nbins: int = 6) -> np.ndarray:
    """n isotropic Gaussian blobs with peak amplitude LogNormal(mu, s) and
    sigma ~ U(sigma_range); sigmas are grouped into nbins bins whose deltas
    are blurred together."""
    vol = np.zeros(shape, dtype=np.float64, order="F")
    idx = [g.integers(0, s, size=n) for s in shape]
    sig = g.uniform(sigma_range[0], sigma_range[1], size=n)
    amp = g.lognormal(lognorm[0], lognorm[1], size=n)
    edges = np.linspace(sigma_range[0], sigma_range[1], nbins + 1)
    b = np.clip(np.digitize(sig, edges) - 1, 0, nbins - 1)
    for k in range(nbins):
        sel = b == k
        if not sel.any():
            continue
        s = 0.5 * (edges[k] + edges[k + 1])
        d = np.zeros(shape, dtype=np.float64, order="F")
        # peak amplitude A <-> delta weight A (2 pi)^(3/2) s^3 under a unit-sum Gaussian
        np.add.at(d, tuple(i[sel] for i in idx), amp[sel] * (2 * np.pi) ** 1.5 * s ** 3)
        if _ndi is not None:
            vol += _ndi.gaussian_filter(d, s, mode="wrap", truncate=4.0)
        else:  # pragma: no cover
            vol += d
    return vol


def smooth_background(shape, g: np.random.Generator, mean: float = 100.0, n: int = 3) -> np.ndarray:
    """mean + sum of n low-frequency sinusoids (|k| <= 2 cycles/volume, a ~ U[5, 20])."""
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) / s for s in shape], indexing="ij")
    bg = np.full(shape, mean, dtype=np.float64)
    for _ in range(n):
        k = g.integers(-2, 3, size=3)
        if not k.any():
            k[0] = 1
        a = g.uniform(5, 20)
        ph = g.uniform(0, 2 * np.pi)
        bg += a * np.sin(2 * np.pi * (k[0] * grids[0] + k[1] * grids[1] + k[2] * grids[2]) + ph)
    return bg


def poisson_u16(lam: np.ndarray, g: np.random.Generator) -> np.ndarray:
    lam = np.clip(lam, 0, None)
    counts = g.poisson(lam)
    return np.asfortranarray(np.clip(counts, 0, 65535).astype(np.uint16))


def emitter_volume(shape, psf, n_emitters=200, amp_range=(500, 3000), background=100.0, config_id=1):
    """Point emitters convolved with the PSF, plus a flat background, Poisson noise."""
    g0, g1 = rng(config_id, 0), rng(config_id, 1)
    obj = point_emitters(shape, n_emitters, amp_range, g0)
    img = _fft_convolve_circular(obj, psf) + background
    return poisson_u16(img, g1)


def blob_volume(shape, psf, n_blobs=20000, config_id=2, bg_mean=100.0):
    """Gaussian blobs + smooth sinusoidal background, blurred by the PSF, Poisson noise."""
    g0, g1, g2 = rng(config_id, 0), rng(config_id, 1), rng(config_id, 2)
    obj = gaussian_blobs(shape, n_blobs, g0) + smooth_background(shape, g1, bg_mean)
    img = _fft_convolve_circular(obj, psf)
    return poisson_u16(img, g2)


def config1():
    """128x128x64 uint16, 200 emitters (amp U[500,3000]), PSF 31x31x41
    (sigma_xy 1.2, sigma_z 3.0), background 100, Poisson."""
    psf = gaussian_psf((31, 31, 41), 1.2, 3.0)
    return emitter_volume((128, 128, 64), psf, 200, (500, 3000), 100.0, 1), psf


def config2(shape=(512, 512, 256), n_blobs=20000):
    """512x512x256 uint16, 20k blobs, smooth background (mean 100), blurred by
    a 64x64x101 PSF (sigma_xy 1.0, sigma_z 2.5, sheared 31.5 deg), Poisson."""
    psf = gaussian_psf((64, 64, 101), 1.0, 2.5, 31.5)
    return blob_volume(shape, psf, n_blobs, 2), psf

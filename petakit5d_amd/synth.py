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


def rng(seed_id: int, stream: int = 0) -> np.random.Generator:
    """Independent seeded sub-stream `stream` of the workload seeded by `seed_id`."""
    return np.random.default_rng([0x5EED0000 + int(seed_id), int(stream)])


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
    pad = _wrapped_kernel(psf, d)
    if _sfft is not None:
        F = _sfft.rfftn(vol, workers=-1) * _sfft.rfftn(pad, workers=-1)
        return _sfft.irfftn(F, s=d, workers=-1)
    F = np.fft.rfftn(vol) * np.fft.rfftn(pad)
    return np.fft.irfftn(F, s=d)


def _wrapped_kernel(psf: np.ndarray, d) -> np.ndarray:
    """PSF placed with its floor(p/2) element at the origin, wrapped modulo the
    volume size d (a PSF larger than the volume folds onto it)."""
    pad = np.zeros(d, dtype=np.float64)
    ix = np.ix_(*[(np.arange(ps) - ps // 2) % n for ps, n in zip(psf.shape, d)])
    np.add.at(pad, ix, psf.astype(np.float64))
    return pad


def point_emitters(shape, n: int, amp_range, g: np.random.Generator) -> np.ndarray:
    vol = np.zeros(shape, dtype=np.float64, order="F")
    idx = [g.integers(0, s, size=n) for s in shape]
    amps = g.uniform(amp_range[0], amp_range[1], size=n)
    np.add.at(vol, tuple(idx), amps)
    return vol


def _blob_params(shape, n, g, sigma_range=(0.5, 2.0), lognorm=(6.0, 0.8), nbins=6):
    idx = [g.integers(0, s, size=n) for s in shape]
    sig = g.uniform(sigma_range[0], sigma_range[1], size=n)
    amp = g.lognormal(lognorm[0], lognorm[1], size=n)
    edges = np.linspace(sigma_range[0], sigma_range[1], nbins + 1)
    b = np.clip(np.digitize(sig, edges) - 1, 0, nbins - 1)
    centers = 0.5 * (edges[:-1] + edges[1:])
    # peak amplitude A <-> delta weight A (2 pi)^(3/2) s^3 under a unit-sum Gaussian
    w = amp * (2 * np.pi) ** 1.5 * centers[b] ** 3
    return idx, b, w, centers


def gaussian_blobs(shape, n: int, g: np.random.Generator, sigma_range=(0.5, 2.0), lognorm=(6.0, 0.8),
                   nbins: int = 6) -> np.ndarray:
    """n isotropic Gaussian blobs with peak amplitude LogNormal(mu, s) and
    sigma ~ U(sigma_range); sigmas are grouped into nbins bins whose deltas
    are blurred together (circular boundary)."""
    idx, b, w, centers = _blob_params(shape, n, g, sigma_range, lognorm, nbins)
    vol = np.zeros(shape, dtype=np.float64, order="F")
    for k in range(nbins):
        sel = b == k
        if not sel.any():
            continue
        d = np.zeros(shape, dtype=np.float64, order="F")
        np.add.at(d, tuple(i[sel] for i in idx), w[sel])
        if _ndi is not None:
            vol += _ndi.gaussian_filter(d, centers[k], mode="wrap", truncate=4.0)
        else:  # pragma: no cover
            vol += d
    return vol


def smooth_background(shape, g: np.random.Generator, mean: float = 100.0, n: int = 3) -> np.ndarray:
    """mean + sum of n low-frequency sinusoids (|k| <= 2 cycles/volume, a ~ U[5, 20])."""
    bg = np.full(shape, mean, dtype=np.float64)
    for a, k, ph in _bg_params(g, n):
        t = [np.arange(s, dtype=np.float64) / s * k[d] for d, s in enumerate(shape)]
        bg += a * np.sin(2 * np.pi * (t[0][:, None, None] + t[1][None, :, None] + t[2][None, None, :]) + ph)
    return bg


def _bg_params(g, n):
    out = []
    for _ in range(n):
        k = g.integers(-2, 3, size=3)
        if not k.any():
            k[0] = 1
        out.append((g.uniform(5, 20), k, g.uniform(0, 2 * np.pi)))
    return out


def poisson_u16(lam: np.ndarray, g: np.random.Generator) -> np.ndarray:
    lam = np.clip(lam, 0, None)
    counts = g.poisson(lam)
    return np.asfortranarray(np.clip(counts, 0, 65535).astype(np.uint16))


def emitter_volume(shape, psf, n_emitters=200, amp_range=(500, 3000), background=100.0, seed_id=1):
    """Point emitters convolved with the PSF, plus a flat background, Poisson noise."""
    g0, g1 = rng(seed_id, 0), rng(seed_id, 1)
    obj = point_emitters(shape, n_emitters, amp_range, g0)
    img = _fft_convolve_circular(obj, psf) + background
    return poisson_u16(img, g1)


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


def _blob_volume_torch(shape, psf, n_blobs, seed_id, bg_mean, dtype=None, as_tensor=False):
    """GPU rendering of blob_volume: blobs, background and PSF blur in one
    Fourier-domain product (float64, or float32 for very large volumes),
    Poisson noise from a seeded torch generator.  Same model as the host
    version; different random stream for the noise.  as_tensor=True returns
    the column-major uint16 CUDA tensor instead of a host array."""
    import torch
    dev = torch.device("cuda")
    fdt = dtype or torch.float64
    g0, g1, g2 = rng(seed_id, 0), rng(seed_id, 1), rng(seed_id, 2)
    idx, b, w, centers = _blob_params(shape, n_blobs, g0)
    n1, n2, n3 = shape
    # torch tensors are indexed [z, y, x] = MATLAB (x=dim1 contiguous) reversed
    flat = torch.from_numpy(idx[0] + n1 * (idx[1] + n2 * idx[2])).to(dev)
    rfft = lambda v: torch.fft.rfftn(v, dim=(0, 1, 2))  # noqa: E731
    acc = None
    for k in range(len(centers)):
        sel = torch.from_numpy(b == k).to(dev)
        if not bool(sel.any()):
            continue
        d = torch.zeros(n1 * n2 * n3, dtype=fdt, device=dev)
        d.index_add_(0, flat[sel], torch.from_numpy(w).to(dev, fdt)[sel])
        s = float(centers[k])
        # circular Gaussian kernel sampled on the grid, unit sum
        ax = [torch.arange(m, dtype=fdt, device=dev) for m in (n3, n2, n1)]
        ax = [torch.minimum(a, m - a) for a, m in zip(ax, (n3, n2, n1))]
        ker = torch.exp(-(ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2)
                        / (2 * s * s))
        ker /= ker.sum()
        term = rfft(d.view(n3, n2, n1)) * rfft(ker)
        acc = term if acc is None else acc + term
        del d, ker
    bg = torch.full((n3, n2, n1), bg_mean, dtype=fdt, device=dev)
    for a, kk, ph in _bg_params(g1, 3):
        t = [torch.arange(m, dtype=fdt, device=dev) / m * kk[dd] for dd, m in enumerate(shape)]
        bg += a * torch.sin(2 * np.pi * (t[2][:, None, None] + t[1][None, :, None] + t[0][None, None, :]) + ph)
    acc = acc + rfft(bg)
    del bg
    p = torch.from_numpy(np.ascontiguousarray(_wrapped_kernel(psf, shape).transpose(2, 1, 0))).to(dev, fdt)
    img = torch.fft.irfftn(acc * rfft(p), s=(n3, n2, n1), dim=(0, 1, 2)).clamp_(min=0)
    del acc, p
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(g2.integers(0, 2 ** 62)))
    counts = torch.poisson(img, generator=gen).clamp_(max=65535)
    del img
    if as_tensor:
        out = counts.to(torch.int32).to(torch.uint16).permute(2, 1, 0)  # column-major view, MATLAB dims
        del counts
        torch.cuda.empty_cache()
        return out
    counts = counts.to(torch.int32).cpu().numpy()
    return np.asfortranarray(counts.astype(np.uint16).transpose(2, 1, 0))


def blob_volume(shape, psf, n_blobs=20000, seed_id=2, bg_mean=100.0, device=None):
    """Gaussian blobs + smooth sinusoidal background, blurred by the PSF,
    Poisson noise.  device='cuda' renders on the GPU (much faster for large
    volumes); the default picks the GPU when one is visible."""
    if device is None:
        device = "cuda" if _gpu_available() else "cpu"
    if device == "cuda":
        return _blob_volume_torch(tuple(int(s) for s in shape), psf, n_blobs, seed_id, bg_mean)
    g0, g1, g2 = rng(seed_id, 0), rng(seed_id, 1), rng(seed_id, 2)
    obj = gaussian_blobs(shape, n_blobs, g0) + smooth_background(shape, g1, bg_mean)
    img = _fft_convolve_circular(obj, psf)
    return poisson_u16(img, g2)


def emitter_case():
    """128x128x64 uint16, 200 emitters (amp U[500,3000]), PSF 31x31x41
    (sigma_xy 1.2, sigma_z 3.0), background 100, Poisson."""
    psf = gaussian_psf((31, 31, 41), 1.2, 3.0)
    return emitter_volume((128, 128, 64), psf, 200, (500, 3000), 100.0, 1), psf


def sheared_psf():
    """64x64x101 Gaussian PSF, sigma_xy 1.0, sigma_z 2.5, sheared 31.5 deg along dim 2."""
    return gaussian_psf((64, 64, 101), 1.0, 2.5, 31.5)


def blob_case(shape=(512, 512, 256), n_blobs=20000, device=None):
    """512x512x256 uint16, 20k blobs, smooth background (mean 100), blurred by
    the sheared 64x64x101 PSF, Poisson."""
    psf = sheared_psf()
    return blob_volume(shape, psf, n_blobs, 2, device=device), psf


def large_blob_case(shape=(2048, 2048, 512), n_blobs=500000):
    """2048x2048x512 uint16 (2^31 voxels), 500k blobs, the sheared 64x64x101 PSF;
    rendered on the GPU in float32 and returned as a column-major CUDA
    tensor (it never leaves HBM)."""
    import torch
    psf = sheared_psf()
    t = _blob_volume_torch(tuple(shape), psf, n_blobs, 5, 100.0, dtype=torch.float32, as_tensor=True)
    return t, psf

"""Writes the committed golden vectors in tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

Nothing here uses the oracle or the product.  Every expected output comes
from an independent computation:
  * fft_*.npz   -- numpy.fft.fftn (pocketfft) in float64;
  * rl_*.npz    -- Richardson-Lucy written with explicit circular
                   convolution / correlation sums over the PSF support (no
                   FFT), following decon_lucy_function.m:116-143 and
                   decon_lucy_omw_function.m:105-112 literally, in float64.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = 2.0 ** -52


def circ_conv(vol, psf):
    """(vol (*) psf)[x] = sum_o psf[o] vol[x - (o - c)], c = floor(p/2): the
    operation real(ifftn(psf2otf(psf) .* fftn(vol))) computes."""
    out = np.zeros_like(vol)
    c = [s // 2 for s in psf.shape]
    for o in np.ndindex(*psf.shape):
        w = psf[o]
        if w == 0:
            continue
        out += w * np.roll(vol, shift=tuple(o[d] - c[d] for d in range(3)), axis=(0, 1, 2))
    return out


def circ_corr(vol, psf):
    """sum_o psf[o] vol[x + (o - c)]: real(ifftn(conj(OTF) .* fftn(vol)))."""
    out = np.zeros_like(vol)
    c = [s // 2 for s in psf.shape]
    for o in np.ndindex(*psf.shape):
        w = psf[o]
        if w == 0:
            continue
        out += w * np.roll(vol, shift=tuple(-(o[d] - c[d]) for d in range(3)), axis=(0, 1, 2))
    return out


def rl_direct(I, psf, numit, background=0.0):
    """decon_lucy_function (CPU branch, float64) with direct sums."""
    psf = psf / psf.sum()
    I = np.maximum(I - background, 0)
    J2 = I.copy()
    J3 = np.zeros_like(I)
    J4 = np.zeros_like(I)
    Y = np.zeros_like(I)
    lam = 0.0
    lams = []
    for k in range(1, numit + 1):
        if k > 2:
            lam = float(np.sum((J2 - Y) * J4) / (np.sum(J4 * J4) + EPS))
            lam = max(min(lam, 1.0), 0.0)
            J4 = J2 - Y
        elif k == 2:
            J4 = J2 - Y
        lams.append(lam)
        Y = np.maximum(J2 + lam * (J2 - J3), 0)
        rb = np.maximum(circ_conv(Y, psf), EPS)
        rb = I / rb
        J3 = J2
        J2 = np.maximum(Y * circ_corr(rb, psf), 0)
    return J2, np.array(lams)


def omw_direct(I, psf_f, psf_b, numit, background=0.0):
    """decon_lucy_omw_function (CPU branch, float64) with direct sums."""
    pf = psf_f / psf_f.sum()
    pb = psf_b / psf_b.sum()
    I = np.maximum(I - background, 0)
    J = I.copy()
    for _ in range(numit):
        cx = np.maximum(circ_conv(J, pf), EPS)
        J = np.maximum(circ_conv(I / cx, pb) * J, 0)
    return J


def main():
    rng = np.random.default_rng(20240601)
    x = rng.standard_normal((6, 10, 15)) + 1j * rng.standard_normal((6, 10, 15))
    np.savez(os.path.join(HERE, "fft_6x10x15.npz"), x=x, X=np.fft.fftn(x))
    xr = rng.standard_normal((8, 9, 25))
    np.savez(os.path.join(HERE, "fft_real_8x9x25.npz"), x=xr, X=np.fft.fftn(xr))

    I = rng.integers(90, 400, size=(8, 6, 5)).astype(np.float64)
    psf = rng.random((3, 3, 3)) + 0.05
    outs = {f"J{k}": rl_direct(I, psf, k, background=100.0)[0] for k in (1, 2, 3, 5)}
    _, lams = rl_direct(I, psf, 5, background=100.0)
    np.savez(os.path.join(HERE, "rl_8x6x5.npz"), I=I, psf=psf, background=100.0, lambdas=lams, **outs)

    psf_f = rng.random((3, 3, 3)) + 0.05
    psf_b = rng.random((3, 3, 1)) + 0.05
    I2 = rng.integers(0, 300, size=(7, 6, 4)).astype(np.float64)
    np.savez(os.path.join(HERE, "omw_7x6x4.npz"), I=I2, psf_f=psf_f, psf_b=psf_b, background=50.0,
             J2=omw_direct(I2, psf_f, psf_b, 2, 50.0), J3=omw_direct(I2, psf_f, psf_b, 3, 50.0))


if __name__ == "__main__":
    main()

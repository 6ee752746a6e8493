"""numpy / scipy restatement of omw_backprojector_generation.

TEST INFRASTRUCTURE ONLY: the checker for the engine's back-projector
generation.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it.

Follows microscopeDataProcessing/decon/omw_backprojector_generation.m:1-166
step by step (line numbers in comments).  MATLAB toolbox semantics restated:
  * strel('sphere', r): voxels with x^2+y^2+z^2 <= r^2;
  * imerode pads with true, imdilate with false; imopen = dilate(erode),
    imclose = erode(dilate);
  * bwareaopen / bwconncomp / bwlabeln: 26-connectivity in 3-D, 8 in 2-D;
    labels numbered by the first voxel in column-major scan order;
  * bwconvhull(BW, 'union'): convex hull of the corners (+-0.5) of all
    foreground pixels; a pixel is set when its centre lies inside or on it;
  * knnsearch K=2: exact Euclidean distances in double, ties broken by the
    lower edge index;
  * cart2sph / sph2cart: MATLAB's formulas (atan2 / hypot, then cos / sin),
    evaluated with the C library's functions.
These toolbox behaviours cannot be executed here (no MATLAB), so parity of
this function with the reference is unpinned beyond the invariants its tests
check.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage as ndi


def ball(r: int) -> np.ndarray:
    g = np.arange(-r, r + 1)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return (x * x + y * y + z * z) <= r * r


def imerode(m, se):
    return ndi.binary_erosion(m, structure=se, border_value=1)


def imdilate(m, se):
    return ndi.binary_dilation(m, structure=se, border_value=0)


def label_colmajor(m, conn_full=True):
    """Connected components labelled in MATLAB order (first voxel in
    column-major order gets label 1).  Returns (labels, count, sizes)."""
    st = np.ones((3,) * m.ndim, bool) if conn_full else None
    lab, n = ndi.label(m, structure=st)
    if n == 0:
        return lab, 0, np.zeros(0, np.int64)
    flat = lab.reshape(-1, order="F")
    idx = np.arange(flat.size)
    first = np.full(n + 1, flat.size, dtype=np.int64)
    nz = flat > 0
    np.minimum.at(first, flat[nz], idx[nz])
    order = np.argsort(first[1:], kind="stable") + 1  # old labels sorted by first voxel
    remap = np.zeros(n + 1, np.int64)
    remap[order] = np.arange(1, n + 1)
    lab = remap[lab]
    sizes = np.bincount(lab.reshape(-1), minlength=n + 1)[1:]
    return lab, n, sizes


def bwareaopen(m, P):
    lab, n, sizes = label_colmajor(m)
    keep = np.concatenate([[False], sizes >= P])
    return keep[lab]


def _hull(points):
    """Monotone-chain convex hull (counter-clockwise, no collinear points)."""
    pts = sorted(set(points))
    if len(pts) <= 2:
        return pts

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


def bwconvhull(img2):
    """bwconvhull(BW, 'union') on a 2-D boolean image (any orientation)."""
    rr, cc = np.nonzero(img2)
    if rr.size == 0:
        return img2.copy()
    pts = []
    for r in np.unique(rr):  # extreme pixels of each row suffice for the hull
        cs = cc[rr == r]
        for c in (cs.min(), cs.max()):
            for dr in (-0.5, 0.5):
                for dc in (-0.5, 0.5):
                    pts.append((float(r + dr), float(c + dc)))
    h = _hull(pts)
    out = np.zeros_like(img2, dtype=bool)
    R, C = np.meshgrid(np.arange(img2.shape[0]), np.arange(img2.shape[1]), indexing="ij")
    inside = np.ones(img2.shape, bool)
    n = len(h)
    for i in range(n):
        (r0, c0), (r1, c1) = h[i], h[(i + 1) % n]
        # counter-clockwise hull: centre inside or on the edge when cross >= 0
        cr = (r1 - r0) * (C - c0) - (c1 - c0) * (R - r0)
        inside &= cr >= -1e-9
    out[inside] = True
    return out


def knn2(edge_pts, query_pts, chunk=4096):
    """Two nearest edge points per query (Euclidean, double), ties -> lower index.

    Candidates come from a k-d tree (scipy cKDTree, 8 nearest); their squared
    distances are recomputed with the same formula as the exhaustive search
    and ranked by (distance, index).  A query whose 8th candidate is not
    clearly farther than its 2nd falls back to the exhaustive search, so the
    result is that of the exhaustive search in every case."""
    E = np.asarray(edge_pts, np.float64)
    Qp = np.asarray(query_pts, np.float64)
    if E.shape[0] <= 8 or Qp.shape[0] < 4096:
        return _knn2_exhaustive(E, Qp, chunk)
    from scipy.spatial import cKDTree
    dk, ik = cKDTree(E).query(Qp, k=8)
    d2 = ((Qp[:, None, 0] - E[ik, 0]) ** 2 + (Qp[:, None, 1] - E[ik, 1]) ** 2 + (Qp[:, None, 2] - E[ik, 2]) ** 2)
    order = np.lexsort((ik, d2), axis=1)[:, :2]
    inds = np.take_along_axis(ik, order, axis=1)
    dsq = np.take_along_axis(d2, order, axis=1)
    # exact ranking is only guaranteed when every point outside the 8 is farther than the 2nd
    unsure = dk[:, 7] <= np.sqrt(dsq[:, 1]) * (1 + 1e-9) + 1e-12
    if unsure.any():
        ui, ud = _knn2_exhaustive(E, Qp[unsure], chunk)
        inds[unsure] = ui
        dsq[unsure] = ud ** 2
        d = np.sqrt(dsq)
        d[unsure] = ud
        return inds, d
    return inds, np.sqrt(dsq)


def _knn2_exhaustive(edge_pts, query_pts, chunk=4096):
    nq = query_pts.shape[0]
    inds = np.zeros((nq, 2), np.int64)
    d = np.zeros((nq, 2), np.float64)
    for s in range(0, nq, chunk):
        q = query_pts[s:s + chunk]
        d2 = ((q[:, None, 0] - edge_pts[None, :, 0]) ** 2 + (q[:, None, 1] - edge_pts[None, :, 1]) ** 2
              + (q[:, None, 2] - edge_pts[None, :, 2]) ** 2)
        o = np.argsort(d2, axis=1, kind="stable")[:, :2]
        inds[s:s + chunk] = o
        d[s:s + chunk] = np.sqrt(np.take_along_axis(d2, o, axis=1))
    return inds, d


def otf_mask_morphology(mask, area_thresh):
    """omw_backprojector_generation.m:44-48 on a thresholded OTF mask: imopen,
    bwareaopen, imclose, imopen, bwareaopen with strel('sphere', 2)."""
    b2 = ball(2)
    mask = imdilate(imerode(mask, b2), b2)  # :44 imopen
    mask = bwareaopen(mask, area_thresh)  # :45
    mask = imerode(imdilate(mask, b2), b2)  # :46 imclose
    mask = imdilate(imerode(mask, b2), b2)  # :47 imopen
    return bwareaopen(mask, area_thresh)  # :48


def skew_detect(mask):
    """omw_backprojector_generation.m:57-71: skewed when the x-z projection has
    more than one component, covers every z, and peaks more than one plane
    from the centre."""
    xz = mask.sum(axis=0) > 0  # squeeze(sum(OTF_mask, 1)) > 0 : (n2, n3)
    _, ncc, _ = label_colmajor(xz)
    if ncc > 1:
        line = xz.sum(axis=0)  # sum(OTF_mask_xz, 1): along z
        pind = int(np.argmax(line)) + 1
        if np.all(line > 0) and abs(pind - (line.size + 1) / 2) > 1:
            return True
    return False


def skewed_components(mask):
    """omw_backprojector_generation.m:73-88: label the mask (bwlabeln, labels
    by first voxel in column-major order); more than three components: keep
    the three largest (maxk) and relabel; one component: it becomes label 2;
    then cat(3, L == 3, L == 2, L == 1) -- right, centre, left along dim 3."""
    L, n, sizes = label_colmajor(mask)
    if n + 1 > 4:
        top = np.argsort(-sizes, kind="stable")[:3] + 1
        mask = np.isin(L, top)
        L, n, sizes = label_colmajor(mask)
    elif n + 1 == 2:
        L = (L + 1) * mask
    return np.concatenate([L == 3, L == 2, L == 1], axis=2)


def largest_component(mask):
    """omw_backprojector_generation.m:90-93: the largest 26-connected
    component (the first in label order among equals)."""
    L, n, sizes = label_colmajor(mask)
    return L == (int(np.argmax(sizes)) + 1)


def slice_hulls(mask):
    """omw_backprojector_generation.m:98-109: bwconvhull of every z-slice, then
    every y-slice, then every x-slice that held mask voxels before the pass
    (index sets taken first)."""
    mask = mask.copy()
    zinds = np.nonzero(mask.sum(axis=(0, 1)))[0]
    xinds = np.nonzero(mask.sum(axis=(0, 2)))[0]
    yinds = np.nonzero(mask.sum(axis=(1, 2)))[0]
    for z in zinds:
        mask[:, :, z] = bwconvhull(mask[:, :, z])
    for y in yinds:
        mask[y, :, :] = bwconvhull(mask[y, :, :])
    for x in xinds:
        mask[:, x, :] = bwconvhull(mask[:, x, :])
    return mask


def edge_radius(E, Q, Rn):
    """omw_backprojector_generation.m:135-139 literally: [inds, d_mat] =
    knnsearch(E, Q, 'K', 2); d_1 = d_mat(inds(:,1), 1); d_2 = d_mat(inds(:,2), 2)
    -- d_mat's rows indexed by EDGE index, as the reference writes it;
    rn = (Rn(inds(:,1)) .* d_2 + Rn(inds(:,2)) .* d_1) ./ (d_1 + d_2)."""
    inds, dmat = knn2(E, Q)  # :135
    d1 = dmat[inds[:, 0], 0]  # :137
    d2 = dmat[inds[:, 1], 1]  # :138
    with np.errstate(invalid="ignore", divide="ignore"):
        return (Rn[inds[:, 0]] * d2 + Rn[inds[:, 1]] * d1) / (d1 + d2)  # :139


def omw_backprojector_generation(psf, alpha, skewed=None, OTFCumThresh=0.9, OTFAreaThresh=100,
                                 hannWinBounds=(0.8, 1.0), minIntThrsh=1e-3):
    """Returns (b_omw, OTF_bp_omw, abs_OTF_c, OTF_mask, skewed_used)."""
    psf = np.asarray(psf)
    if psf.size == 0 or not np.any(psf):  # :25-27
        raise ValueError("A valid nonzero PSF needs to be provided!")
    psf = psf.astype(np.float32)
    psf = (psf / np.float32(psf.astype(np.float64).sum())).astype(np.float32)  # :29-30
    psz = psf.shape
    # OTF = decon_psf2otf(psf) at the PSF size (:32), single precision result
    sh = np.roll(psf.astype(np.float64), shift=tuple(-(s // 2) for s in psz), axis=(0, 1, 2))
    OTF = np.fft.fftn(sh).astype(np.complex64)
    abs_OTF = np.abs(OTF).astype(np.float32)  # :33
    abs_OTF_c = np.fft.fftshift(abs_OTF)  # :34
    vals = np.sort(abs_OTF.astype(np.float64).reshape(-1, order="F"))[::-1]  # :35
    cs = np.cumsum(vals)
    tind = int(np.argmax(cs > cs[-1] * OTFCumThresh))  # :38 (sequential sum)
    dc = np.float32(np.float32(vals[tind]) / abs_OTF[0, 0, 0])  # :39 (single)
    dc = max(dc, np.float32(minIntThrsh))  # :41
    mask = otf_mask_morphology(abs_OTF_c > abs_OTF[0, 0, 0] * dc, OTFAreaThresh)  # :43-48
    if not mask.any():  # :50-52
        raise ValueError("The OTF mask is empty, check the PSF and OTF-related parameters!")

    if skewed is None:  # :57-71
        skewed = skew_detect(mask)
    mask = skewed_components(mask) if skewed else largest_component(mask)  # :73-94
    psz3 = mask.shape
    mask = slice_hulls(mask)  # :98-109

    edge = mask & ~imerode(mask, ball(1))  # :122
    c = (np.array(psz3, np.float64) + 1) / 2  # :123
    Y, X, Z = np.meshgrid(np.arange(1, psz3[0] + 1) - c[0], np.arange(1, psz3[1] + 1) - c[1],
                          np.arange(1, psz3[2] + 1) - c[2], indexing="ij")
    # cart2sph / sph2cart (:129-132) with the C library's atan2, hypot, cos and
    # sin, so exact-distance ties between lattice directions resolve the same
    # way as in any other glibc-based implementation.  atan2/cos/sin come from
    # the math module (thin libm wrappers); math.hypot is CPython's own
    # algorithm and differs from libm's in the last bit, so hypot is taken
    # from libm directly.
    import ctypes
    import math
    _libm = ctypes.CDLL("libm.so.6")
    _libm.hypot.restype = ctypes.c_double
    _libm.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    hypot = _libm.hypot
    shp = X.shape
    Xf, Yf, Zf = X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
    rs = np.empty(Xf.size)
    Xn = np.empty(Xf.size)
    Yn = np.empty(Xf.size)
    Zn = np.empty(Xf.size)
    for i in range(Xf.size):
        x, y, z = float(Xf[i]), float(Yf[i]), float(Zf[i])
        hxy = hypot(x, y)
        rs[i] = hypot(hxy, z)
        el, az = math.atan2(z, hxy), math.atan2(y, x)
        rc = math.cos(el)
        Xn[i], Yn[i], Zn[i] = rc * math.cos(az), rc * math.sin(az), math.sin(el)
    rs, Xn, Yn, Zn = rs.reshape(shp), Xn.reshape(shp), Yn.reshape(shp), Zn.reshape(shp)
    f = lambda a: a.reshape(-1, order="F")  # noqa: E731
    ef = f(edge)
    E = np.stack([f(Yn)[ef], f(Xn)[ef], f(Zn)[ef]], axis=1)
    Q = np.stack([f(Yn), f(Xn), f(Zn)], axis=1)
    rn = edge_radius(E, Q, f(rs)[ef])  # :135-139
    with np.errstate(invalid="ignore", divide="ignore"):
        bw = f(rs) / rn  # :141
    lo, hi = float(hannWinBounds[0]), float(hannWinBounds[1])
    with np.errstate(invalid="ignore"):
        win = (bw < lo).astype(np.float64) + np.cos(np.pi * (bw - lo) / 2 / (hi - lo)) ** 2 * ((bw >= lo) & (bw < hi))
    m = win.reshape(psz3, order="F")  # :148
    if skewed:  # :150-153
        m = m.reshape(psz[0], psz[1], psz[2], 3, order="F").sum(axis=3)
        mask = mask.reshape(psz[0], psz[1], psz[2], 3, order="F").sum(axis=3)
    om = np.fft.ifftshift(m)  # :155
    den = (abs_OTF * abs_OTF + np.float32(alpha)).astype(np.float32)  # :158, single
    w = (OTF.real / den + 1j * (-OTF.imag / den)).astype(np.complex64)
    omw = (om.astype(np.float32) * w).astype(np.complex64)  # :161
    b = np.fft.fftshift(np.real(np.fft.ifftn(omw.astype(np.complex128)))).astype(np.float32)  # :163
    return b, omw, abs_OTF_c, mask, bool(skewed)


# ---------------------------------------------------------------------------------------------------------------------
# BackProjector (third_parties/WBDeconvolution/BackProjector.m:1-233): restated with numpy.  The reference's
# PSF is single, so its FWHMs, cutoffs and Butterworth masks are single-precision values (float32 here); its
# spectra are double (complex128 here).
def wb_fwhm(y):
    """fwhm(x, y), x = 1:N (BackProjector.m:474-509): NaN for a step-like profile; raises IndexError where the
    reference's unbounded crossing search would index past the end."""
    y = np.asarray(y, np.float32)
    y = (y / y.max()).astype(np.float32)
    N = y.size
    lev = np.float32(0.5)
    center = int(np.argmax(y)) if y[0] < lev else int(np.argmin(y))  # 0-based, first occurrence
    s = np.sign(y - lev)
    i = 1
    while s[i] == s[i - 1]:
        i += 1
        if i >= N:
            raise IndexError("index exceeds the profile (leading crossing)")
    tlead = np.float32(i) + (lev - y[i - 1]) / (y[i] - y[i - 1])
    i = center + 1
    if i > N - 1:
        raise IndexError("index exceeds the profile (trailing crossing)")
    while i <= N - 2 and s[i] == s[i - 1]:
        i += 1
    if i != N - 1:
        ttrail = np.float32(i) + (lev - y[i - 1]) / (y[i] - y[i - 1])
        return np.float32(ttrail - tlead)
    return np.float32(np.nan)


def wb_fwhm_psf(p):
    """fwhm_PSF(PSF) with cFlag 0, fitFlag 0 (:257-379): profiles through the first maximum (column-major)."""
    p = np.asarray(p, np.float32)
    if p.ndim == 2:
        p = p[:, :, None]
    sx, sy, sz = p.shape
    if sx == 1 or sy == 1:
        return np.float32(wb_fwhm(p.reshape(-1, order="F"))), np.float32(0), np.float32(0)
    ind = int(np.argmax(p.reshape(-1, order="F")))
    ix, iy, iz = np.unravel_index(ind, p.shape, order="F")
    fx = wb_fwhm(p[:, iy, iz])
    fy = wb_fwhm(p[ix, :, iz])
    fz = wb_fwhm(p[ix, iy, :]) if sz > 1 else np.float32(0)
    return fx, fy, fz


def _mround(x):
    return float(np.sign(x) * np.floor(np.abs(x) + 0.5))


def _mmax(x, y):
    """MATLAB max(x, y) of scalars: NaN is skipped."""
    return y if np.isnan(x) else max(x, y)


def _mmin(x, y):
    """MATLAB min(x, y) of scalars: NaN is skipped."""
    return y if np.isnan(x) else min(x, y)


def wb_backprojector(psf, bp_type="traditional", alpha=0.001, beta=1.0, n=10, resFlag=1, iRes=(0, 0, 0)):
    """Returns (PSF_bp, OTF_bp, info) like BackProjector(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes)."""
    p = np.asarray(psf, np.float32)
    if p.ndim == 2:
        p = p[:, :, None]
    if p.shape[0] == 1 or p.shape[1] == 1:
        raise ValueError("BackProjector needs a 2-D or 3-D PSF")
    S = np.array(p.shape)
    Sc = (S + 1) / 2.0  # :68-70, double
    Soz = int(_mround((S[2] + 1) / 2.0))
    fw = np.array(wb_fwhm_psf(p), np.float32)
    flipped = p[::-1, ::-1, ::-1].astype(np.float64)  # flipPSF (:535-544), double
    otf_flip = np.fft.fftn(np.fft.ifftshift(flipped))
    oabs = np.fft.fftshift(np.abs(otf_flip))
    M = oabs.max()
    on = oabs / M
    if resFlag == 0:
        res = (fw / np.float32(np.sqrt(2.0))).astype(np.float32)
    elif resFlag == 1:
        res = fw
    elif resFlag == 2:
        res = np.asarray(iRes, np.float64)
    else:
        raise ValueError("Processing terminated, please set resFlag as 0, 1, or 2")
    if resFlag == 2:
        t = 1.0 / res / (1.0 / S)
        lo = [int(_mmax(_mround(Sc[a] - t[a]), 1)) for a in range(3)]
        hi = [int(_mmin(_mround(Sc[a] + t[a]), S[a])) for a in range(3)]
    else:
        px = (1.0 / S).astype(np.float32)
        t = ((np.float32(1) / res) / px).astype(np.float32)
        lo = [int(_mmax(_mround(float(np.float32(Sc[a]) - t[a])), 1)) for a in range(3)]
        hi = [int(_mmin(_mround(float(np.float32(Sc[a]) + t[a])), S[a])) for a in range(3)]
    lx = on.max(axis=(1, 2))  # max over z then over y (:117-121)
    ly = on.max(axis=(0, 2))
    lz = on.max(axis=(0, 1))
    # 2-D PSF: squeeze(max(OTF_abs_norm,[],1)) stays a 1 x Sy row; both z indices are 1
    bz = (lz[lo[2] - 1] + lz[hi[2] - 1]) / 2 if S[2] > 1 else on[:, 0, 0].max()
    beta_fp = ((lx[lo[0] - 1] + lx[hi[0] - 1]) / 2 + (ly[lo[1] - 1] + ly[hi[1] - 1]) / 2 + bz) / 3
    if alpha == 1:
        alpha = beta_fp
    if beta == 1:
        beta = beta_fp
    # w = sum(((index - Sc) / t)^2), single unless iRes was given (:167-177)
    I, J, K = np.meshgrid(np.arange(1, S[0] + 1), np.arange(1, S[1] + 1), np.arange(1, S[2] + 1), indexing="ij")
    if resFlag == 2:
        w = (((I - Sc[0]) / t[0]) ** 2 + ((J - Sc[1]) / t[1]) ** 2 + ((K - Sc[2]) / t[2]) ** 2).astype(np.float32)
    else:
        a = (I - Sc[0]).astype(np.float32) / t[0]
        b = (J - Sc[1]).astype(np.float32) / t[1]
        c = (K - Sc[2]).astype(np.float32) / t[2]
        w = (a * a + b * b + c * c).astype(np.float32)

    def butter(ee):
        m = np.float32(1) / np.sqrt(np.float32(1) + np.float32(ee) * np.power(w, np.float32(n)))
        return np.fft.ifftshift(m.astype(np.float64))

    beta_w = 0.0
    if bp_type == "traditional":
        out, otf = flipped, otf_flip
    elif bp_type == "gaussian":  # gen_gaussianPSF_3D (:511-533)
        sig = (fw / np.float32(2.3548)).astype(np.float32)
        q = (sig * sig * np.float32(2)).astype(np.float32)
        coef = np.float32(1) / (np.float32((2 * np.pi) ** 1.5) * sig[0] * sig[1] * sig[2])
        a = (I - Sc[0]).astype(np.float32)
        b = (J - Sc[1]).astype(np.float32)
        c = (K - Sc[2]).astype(np.float32)
        dd = a * a / q[0] + b * b / q[1] + c * c / q[2]
        out = (coef * np.exp(-dd)).astype(np.float32).astype(np.float64)
        otf = np.fft.fftn(np.fft.ifftshift(out))
    elif bp_type == "butterworth":
        otf = butter(1.0 / beta ** 2 - 1.0).astype(np.complex128)
        out = np.fft.fftshift(np.real(np.fft.ifftn(otf)))
    elif bp_type in ("wiener", "wiener-butterworth"):
        onrm = otf_flip / M
        wn = onrm / (np.abs(onrm) ** 2 + alpha)
        if bp_type == "wiener":
            otf = wn
        else:
            wa = np.fft.fftshift(np.abs(wn))
            tl = np.abs(wa[:, :, Soz - 1]).max(axis=1)
            beta_w = (tl[lo[0] - 1] + tl[hi[0] - 1]) / 2
            otf = butter(beta_w / beta ** 2 - 1.0) * wn
        out = np.fft.fftshift(np.real(np.fft.ifftn(otf)))
    else:
        raise ValueError("bp_type does not match any back-projector type")
    info = {"FWHM": [float(x) for x in fw], "cutoff": [float(x) for x in t], "beta_fp": float(beta_fp),
            "alpha": float(alpha), "beta": float(beta), "beta_wienerx": float(beta_w)}
    return out, otf, info

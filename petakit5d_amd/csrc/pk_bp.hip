// pk_bp.hip -- OMW back-projector generation
// (microscopeDataProcessing/decon/omw_backprojector_generation.m:1-166).
//
// The per-voxel nearest-edge search (knnsearch K=2 over every voxel of the
// PSF-sized OTF domain, :135) runs on the GPU as an exact FP64 brute-force
// kernel.  The small sequential toolbox steps -- OTF at the PSF size (any
// length, e.g. a prime 101-slice PSF), sphere-SE morphology, 26-connected
// labelling, per-slice convex hulls -- run on the host in this file.
//
// MATLAB semantics restated (steps follow the reference line by line):
//   strel('sphere', r) = voxels with x^2+y^2+z^2 <= r^2; imerode pads with
//   true, imdilate with false; bwareaopen / bwconncomp / bwlabeln use
//   26-connectivity (8 in 2-D) and number components by their first voxel in
//   column-major order; bwconvhull(BW,'union') = convex hull of the corners of
//   the foreground pixels, pixel set when its centre is inside or on it;
//   knnsearch ties go to the lower edge index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "pk_engine.hpp"

// every product and sum here rounds separately, as MATLAB's elementwise
// single/double arithmetic does (no fused multiply-add contraction)
#pragma clang fp contract(off)

namespace pk {

using cd = std::complex<double>;

// ------------------------------------------------------------------ host FFT (any length)
// Recursive mixed-radix decimation in time; twiddles from a table of the
// line length's roots of unity (double).
static void fft_rec(const cd* x, int s, cd* y, int n, int N, const std::vector<cd>& W, std::vector<cd>& tmp) {
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int p = 2;
    while (n % p) ++p;
    const int m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(x + (size_t)r * s, s * p, y + (size_t)r * m, m, N, W, tmp);
    const int step = N / n;
    cd t[64];
    std::vector<cd> big;
    cd* tt = t;
    if (p > 64) {
        big.resize(p);
        tt = big.data();
    }
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) tt[r] = y[(size_t)r * m + k] * W[(size_t)((int64_t)r * k * step % N)];
        for (int q = 0; q < p; ++q) {
            cd acc = 0;
            for (int r = 0; r < p; ++r) acc += tt[r] * W[(size_t)(((int64_t)r * q % p) * (N / p))];
            tmp[(size_t)q * m + k] = acc;
        }
    }
    for (int i = 0; i < n; ++i) y[i] = tmp[i];
}

// In-place 3-D DFT of a column-major complex volume; inverse includes 1/N.
static void fft3_host(std::vector<cd>& a, const int64_t d[3], bool inverse) {
    const int64_t n1 = d[0], n2 = d[1], n3 = d[2];
    for (int ax = 0; ax < 3; ++ax) {
        const int n = (int)d[ax];
        if (n == 1) continue;
        std::vector<cd> W(n), line(n), out(n), tmp(n);
        for (int e = 0; e < n; ++e) {
            const long double ang = 2.0L * 3.14159265358979323846264338327950288L * e / n;
            W[e] = cd((double)cosl(ang), (double)(inverse ? sinl(ang) : -sinl(ang)));
        }
        const int64_t stride = ax == 0 ? 1 : (ax == 1 ? n1 : n1 * n2);
        const int64_t nlines = n1 * n2 * n3 / n;
        for (int64_t l = 0; l < nlines; ++l) {
            int64_t base;
            if (ax == 0) base = l * n1;
            else if (ax == 1) base = (l % n1) + (l / n1) * n1 * n2;
            else base = l;
            for (int i = 0; i < n; ++i) line[i] = a[(size_t)(base + i * stride)];
            fft_rec(line.data(), 1, out.data(), n, n, W, tmp);
            for (int i = 0; i < n; ++i) a[(size_t)(base + i * stride)] = out[i];
        }
    }
    if (inverse) {
        const double sc = 1.0 / (double)(n1 * n2 * n3);
        for (auto& v : a) v *= sc;
    }
}

// ------------------------------------------------------------------ volume helpers
struct Vol3 {
    int64_t d[3];
    int64_t n() const { return d[0] * d[1] * d[2]; }
    int64_t at(int64_t i, int64_t j, int64_t k) const { return i + d[0] * (j + d[1] * k); }
};

template <class T>
static std::vector<T> shift3(const std::vector<T>& a, const Vol3& v, const int64_t sh[3]) {
    // out(i) = a(i - sh) (circular): fftshift uses sh = floor(n/2), ifftshift -floor(n/2)
    std::vector<T> o(a.size());
    for (int64_t k = 0; k < v.d[2]; ++k)
        for (int64_t j = 0; j < v.d[1]; ++j)
            for (int64_t i = 0; i < v.d[0]; ++i) {
                const int64_t ii = ((i + sh[0]) % v.d[0] + v.d[0]) % v.d[0];
                const int64_t jj = ((j + sh[1]) % v.d[1] + v.d[1]) % v.d[1];
                const int64_t kk = ((k + sh[2]) % v.d[2] + v.d[2]) % v.d[2];
                o[(size_t)v.at(ii, jj, kk)] = a[(size_t)v.at(i, j, k)];
            }
    return o;
}

static std::vector<uint8_t> morph(const std::vector<uint8_t>& m, const Vol3& v, int r, bool erode) {
    std::vector<int> oi, oj, ok;
    for (int k = -r; k <= r; ++k)
        for (int j = -r; j <= r; ++j)
            for (int i = -r; i <= r; ++i)
                if (i * i + j * j + k * k <= r * r) {
                    oi.push_back(i);
                    oj.push_back(j);
                    ok.push_back(k);
                }
    std::vector<uint8_t> o(m.size());
    const int ns = (int)oi.size();
    for (int64_t k = 0; k < v.d[2]; ++k)
        for (int64_t j = 0; j < v.d[1]; ++j)
            for (int64_t i = 0; i < v.d[0]; ++i) {
                bool res = erode;
                for (int s = 0; s < ns; ++s) {
                    const int64_t ii = i + oi[s], jj = j + oj[s], kk = k + ok[s];
                    const bool out = ii < 0 || jj < 0 || kk < 0 || ii >= v.d[0] || jj >= v.d[1] || kk >= v.d[2];
                    const bool val = out ? erode : m[(size_t)v.at(ii, jj, kk)] != 0;  // erode pads true
                    if (erode && !val) {
                        res = false;
                        break;
                    }
                    if (!erode && val) {
                        res = true;
                        break;
                    }
                }
                o[(size_t)v.at(i, j, k)] = res;
            }
    return o;
}
static std::vector<uint8_t> imerode(const std::vector<uint8_t>& m, const Vol3& v, int r) { return morph(m, v, r, true); }
static std::vector<uint8_t> imdilate(const std::vector<uint8_t>& m, const Vol3& v, int r) { return morph(m, v, r, false); }

// Connected components (26-connectivity in 3-D; 8 in 2-D when d[2] == 1),
// labels 1..n in order of each component's first voxel in column-major scan.
static int label_cc(const std::vector<uint8_t>& m, const Vol3& v, std::vector<int>& lab, std::vector<int64_t>& sizes) {
    lab.assign(m.size(), 0);
    sizes.clear();
    int n = 0;
    std::vector<int64_t> stack;
    for (int64_t s = 0; s < v.n(); ++s) {
        if (!m[(size_t)s] || lab[(size_t)s]) continue;
        ++n;
        int64_t cnt = 0;
        stack.push_back(s);
        lab[(size_t)s] = n;
        while (!stack.empty()) {
            const int64_t c = stack.back();
            stack.pop_back();
            ++cnt;
            const int64_t i = c % v.d[0], j = (c / v.d[0]) % v.d[1], k = c / (v.d[0] * v.d[1]);
            for (int dk = -1; dk <= 1; ++dk)
                for (int dj = -1; dj <= 1; ++dj)
                    for (int di = -1; di <= 1; ++di) {
                        const int64_t ii = i + di, jj = j + dj, kk = k + dk;
                        if (ii < 0 || jj < 0 || kk < 0 || ii >= v.d[0] || jj >= v.d[1] || kk >= v.d[2]) continue;
                        const int64_t q = v.at(ii, jj, kk);
                        if (m[(size_t)q] && !lab[(size_t)q]) {
                            lab[(size_t)q] = n;
                            stack.push_back(q);
                        }
                    }
        }
        sizes.push_back(cnt);
    }
    return n;
}

static void bwareaopen(std::vector<uint8_t>& m, const Vol3& v, double P) {
    std::vector<int> lab;
    std::vector<int64_t> sizes;
    label_cc(m, v, lab, sizes);
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] && (double)sizes[(size_t)lab[i] - 1] < P) m[i] = 0;
}

// bwconvhull(BW, 'union') of a 2-D image given by an accessor over [0,R) x [0,C).
template <class Get, class Set>
static void convhull_fill(int R, int C, Get get, Set set) {
    std::vector<std::pair<double, double>> pts;
    for (int r = 0; r < R; ++r) {
        int cmin = -1, cmax = -1;
        for (int c = 0; c < C; ++c)
            if (get(r, c)) {
                if (cmin < 0) cmin = c;
                cmax = c;
            }
        if (cmin < 0) continue;
        for (int c : {cmin, cmax})
            for (double dr : {-0.5, 0.5})
                for (double dc : {-0.5, 0.5}) pts.emplace_back(r + dr, c + dc);
    }
    if (pts.empty()) return;
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    auto cross = [](const std::pair<double, double>& o, const std::pair<double, double>& a,
                    const std::pair<double, double>& b) {
        return (a.first - o.first) * (b.second - o.second) - (a.second - o.second) * (b.first - o.first);
    };
    std::vector<std::pair<double, double>> h;
    if (pts.size() <= 2) {
        h = pts;
    } else {
        std::vector<std::pair<double, double>> lo, up;
        for (auto& p : pts) {
            while (lo.size() >= 2 && cross(lo[lo.size() - 2], lo.back(), p) <= 0) lo.pop_back();
            lo.push_back(p);
        }
        for (auto it = pts.rbegin(); it != pts.rend(); ++it) {
            while (up.size() >= 2 && cross(up[up.size() - 2], up.back(), *it) <= 0) up.pop_back();
            up.push_back(*it);
        }
        lo.pop_back();
        up.pop_back();
        h = lo;
        h.insert(h.end(), up.begin(), up.end());
    }
    const size_t n = h.size();
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            bool in = true;
            for (size_t e = 0; e < n && in; ++e) {
                const auto& a = h[e];
                const auto& b = h[(e + 1) % n];
                const double cr = (b.first - a.first) * (c - a.second) - (b.second - a.second) * (r - a.first);
                in = cr >= -1e-9;
            }
            set(r, c, in);
        }
}

// ------------------------------------------------------------------ GPU nearest edges
constexpr int kKnnTile = 1024;

// For every query point, the two nearest edge points (exact FP64 Euclidean
// distance, non-fused operations so ties resolve identically on any IEEE
// implementation; equal distances keep the lower edge index first).
__global__ void k_knn2(const double* __restrict__ E, int nE, const double* __restrict__ Q, int64_t nQ,
                       int* __restrict__ inds, double* __restrict__ dist) {
    __shared__ double se[3 * kKnnTile];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = q < nQ;
    const double qy = live ? Q[3 * q] : 0.0, qx = live ? Q[3 * q + 1] : 0.0, qz = live ? Q[3 * q + 2] : 0.0;
    double b1 = 1e300, b2 = 1e300;
    int i1 = 0, i2 = 0;
    for (int t0 = 0; t0 < nE; t0 += kKnnTile) {
        const int nt = min(kKnnTile, nE - t0);
        __syncthreads();
        for (int i = threadIdx.x; i < 3 * nt; i += blockDim.x) se[i] = E[3 * (int64_t)t0 + i];
        __syncthreads();
        for (int j = 0; j < nt; ++j) {
            // Separately rounded products and sums, (dy^2 + dx^2) + dz^2, written
            // as plain operators so the file's contract(off) applies (the
            // __dmul_rn family are inline header functions that the default
            // -ffp-contract=fast would fuse): lattice directions produce many
            // distances that tie up to the last bit, and an FMA reorders them.
            const double dy = qy - se[3 * j];
            const double dx = qx - se[3 * j + 1];
            const double dz = qz - se[3 * j + 2];
            const double d2 = (dy * dy + dx * dx) + dz * dz;
            if (d2 < b1) {
                b2 = b1;
                i2 = i1;
                b1 = d2;
                i1 = t0 + j;
            } else if (d2 < b2) {
                b2 = d2;
                i2 = t0 + j;
            }
        }
    }
    if (live) {
        inds[2 * q] = i1;
        inds[2 * q + 1] = i2;
        dist[2 * q] = __dsqrt_rn(b1);
        dist[2 * q + 1] = __dsqrt_rn(b2);
    }
}

// ------------------------------------------------------------------ driver
void pk_ctx_impl::omw_backprojector(const float* psf_in, const int64_t pd[3], double alpha, int skewed_in,
                                    const double thr[5], float* b_out, float* otf_bp_out, float* abs_otf_c_out,
                                    float* mask_out, int* skewed_used) {
    const double otf_thresh = thr[0], area_thresh = thr[1], lo = thr[2], hi = thr[3], min_int = thr[4];
    Vol3 v{{pd[0], pd[1], pd[2]}};
    const int64_t P = v.n();
    if (P == 0) fail(PK_ERR_EMPTY_PSF, "A valid nonzero PSF needs to be provided!");
    bool allzero = true;
    for (int64_t i = 0; i < P && allzero; ++i) allzero = psf_in[i] == 0.f;
    if (allzero) fail(PK_ERR_EMPTY_PSF, "A valid nonzero PSF needs to be provided!");  // :25-27

    // psf = single(psf) ./ sum(psf, 'all') (:29-30)
    double s = 0;
    for (int64_t i = 0; i < P; ++i) s += psf_in[i];
    const float sf = (float)s;
    // OTF = decon_psf2otf(psf) at the PSF size, single result (:32)
    std::vector<cd> a((size_t)P);
    const int64_t h[3] = {pd[0] / 2, pd[1] / 2, pd[2] / 2};
    for (int64_t k = 0; k < pd[2]; ++k)
        for (int64_t j = 0; j < pd[1]; ++j)
            for (int64_t i = 0; i < pd[0]; ++i) {
                const int64_t ii = ((i - h[0]) % pd[0] + pd[0]) % pd[0];
                const int64_t jj = ((j - h[1]) % pd[1] + pd[1]) % pd[1];
                const int64_t kk = ((k - h[2]) % pd[2] + pd[2]) % pd[2];
                a[(size_t)v.at(ii, jj, kk)] = cd((double)(psf_in[v.at(i, j, k)] / sf), 0.0);
            }
    fft3_host(a, pd, false);
    std::vector<std::complex<float>> OTF((size_t)P);
    std::vector<float> absO((size_t)P);
    for (int64_t i = 0; i < P; ++i) {
        OTF[(size_t)i] = std::complex<float>((float)a[(size_t)i].real(), (float)a[(size_t)i].imag());
        absO[(size_t)i] = std::abs(OTF[(size_t)i]);  // :33 (single)
    }
    const int64_t fsh[3] = {pd[0] / 2, pd[1] / 2, pd[2] / 2};
    std::vector<float> absOc = shift3(absO, v, fsh);  // :34 fftshift

    // threshold from the cumulative sorted |OTF| (:35-41)
    std::vector<double> vals(absO.begin(), absO.end());
    std::sort(vals.begin(), vals.end(), std::greater<double>());
    double tot = 0;
    for (double x : vals) tot += x;
    double cs = 0;
    size_t tind = 0;
    for (size_t i = 0; i < vals.size(); ++i) {
        cs += vals[i];
        if (cs > tot * otf_thresh) {
            tind = i;
            break;
        }
    }
    float dc = (float)vals[tind] / absO[0];
    dc = std::max(dc, (float)min_int);
    std::vector<uint8_t> mask((size_t)P);
    const float cut = absO[0] * dc;
    for (int64_t i = 0; i < P; ++i) mask[(size_t)i] = absOc[(size_t)i] > cut;  // :43
    mask = imdilate(imerode(mask, v, 2), v, 2);  // :44 imopen
    bwareaopen(mask, v, area_thresh);            // :45
    mask = imerode(imdilate(mask, v, 2), v, 2);  // :46 imclose
    mask = imdilate(imerode(mask, v, 2), v, 2);  // :47 imopen
    bwareaopen(mask, v, area_thresh);            // :48
    bool any = false;
    for (auto x : mask) any |= x != 0;
    if (!any) fail(PK_ERR_EMPTY_MASK, "The OTF mask is empty, check the PSF and OTF-related parameters!");  // :50-52

    // skewed-space auto detection (:57-71)
    bool skewed = skewed_in > 0;
    if (skewed_in < 0) {
        skewed = false;
        Vol3 v2{{pd[1], pd[2], 1}};
        std::vector<uint8_t> xz((size_t)(pd[1] * pd[2]), 0);
        for (int64_t k = 0; k < pd[2]; ++k)
            for (int64_t j = 0; j < pd[1]; ++j)
                for (int64_t i = 0; i < pd[0]; ++i)
                    if (mask[(size_t)v.at(i, j, k)]) xz[(size_t)(j + pd[1] * k)] = 1;
        std::vector<int> lab;
        std::vector<int64_t> sizes;
        if (label_cc(xz, v2, lab, sizes) > 1) {
            std::vector<int64_t> line((size_t)pd[2], 0);
            for (int64_t k = 0; k < pd[2]; ++k)
                for (int64_t j = 0; j < pd[1]; ++j) line[(size_t)k] += xz[(size_t)(j + pd[1] * k)];
            const int64_t pind = (int64_t)(std::max_element(line.begin(), line.end()) - line.begin()) + 1;
            bool allpos = true;
            for (auto x : line) allpos &= x > 0;
            if (allpos && std::fabs((double)pind - (double)(line.size() + 1) / 2.0) > 1) skewed = true;
        }
    }

    Vol3 w = v;  // working domain (3x along z when skewed)
    if (skewed) {  // :73-88
        std::vector<int> lab;
        std::vector<int64_t> sizes;
        int n = label_cc(mask, v, lab, sizes);
        if (n + 1 > 4) {
            std::vector<int> order((size_t)n);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sizes[(size_t)x] > sizes[(size_t)y]; });
            std::vector<uint8_t> keep((size_t)n + 1, 0);
            for (int t = 0; t < 3; ++t) keep[(size_t)order[(size_t)t] + 1] = 1;
            for (size_t i = 0; i < mask.size(); ++i) mask[i] = keep[(size_t)lab[i]];
            n = label_cc(mask, v, lab, sizes);
        } else if (n + 1 == 2) {
            for (size_t i = 0; i < mask.size(); ++i) lab[i] = (lab[i] + 1) * (mask[i] != 0);
        }
        w = Vol3{{pd[0], pd[1], 3 * pd[2]}};
        std::vector<uint8_t> m3((size_t)w.n(), 0);
        const int want[3] = {3, 2, 1};  // cat(3, L==3, L==2, L==1)
        for (int b = 0; b < 3; ++b)
            for (int64_t i = 0; i < P; ++i) m3[(size_t)(i + b * P)] = lab[(size_t)i] == want[b];
        mask.swap(m3);
    } else {  // :90-93
        std::vector<int> lab;
        std::vector<int64_t> sizes;
        label_cc(mask, v, lab, sizes);
        const int best = (int)(std::max_element(sizes.begin(), sizes.end()) - sizes.begin()) + 1;
        for (size_t i = 0; i < mask.size(); ++i) mask[i] = lab[i] == best;
    }

    // convex hulls per slice along z, y, x (:98-109); index sets taken first
    std::vector<int64_t> zi, yi, xi;
    {
        std::vector<uint8_t> hz((size_t)w.d[2], 0), hy((size_t)w.d[0], 0), hx((size_t)w.d[1], 0);
        for (int64_t k = 0; k < w.d[2]; ++k)
            for (int64_t j = 0; j < w.d[1]; ++j)
                for (int64_t i = 0; i < w.d[0]; ++i)
                    if (mask[(size_t)w.at(i, j, k)]) hz[(size_t)k] = hy[(size_t)i] = hx[(size_t)j] = 1;
        for (int64_t k = 0; k < w.d[2]; ++k) if (hz[(size_t)k]) zi.push_back(k);
        for (int64_t i = 0; i < w.d[0]; ++i) if (hy[(size_t)i]) yi.push_back(i);
        for (int64_t j = 0; j < w.d[1]; ++j) if (hx[(size_t)j]) xi.push_back(j);
    }
    for (int64_t z : zi)
        convhull_fill((int)w.d[0], (int)w.d[1], [&](int r, int c) { return mask[(size_t)w.at(r, c, z)] != 0; },
                      [&](int r, int c, bool b) { mask[(size_t)w.at(r, c, z)] = b; });
    for (int64_t y : yi)
        convhull_fill((int)w.d[1], (int)w.d[2], [&](int r, int c) { return mask[(size_t)w.at(y, r, c)] != 0; },
                      [&](int r, int c, bool b) { mask[(size_t)w.at(y, r, c)] = b; });
    for (int64_t x : xi)
        convhull_fill((int)w.d[0], (int)w.d[2], [&](int r, int c) { return mask[(size_t)w.at(r, x, c)] != 0; },
                      [&](int r, int c, bool b) { mask[(size_t)w.at(r, x, c)] = b; });

    // edge voxels and unit directions (:122-132)
    std::vector<uint8_t> er = imerode(mask, w, 1);
    const int64_t nW = w.n();
    const double c1 = (w.d[0] + 1) / 2.0, c2 = (w.d[1] + 1) / 2.0, c3 = (w.d[2] + 1) / 2.0;
    std::vector<double> Q((size_t)(3 * nW)), rs((size_t)nW);
    std::vector<double> E;
    std::vector<double> Rn;
    for (int64_t k = 0; k < w.d[2]; ++k)
        for (int64_t j = 0; j < w.d[1]; ++j)
            for (int64_t i = 0; i < w.d[0]; ++i) {
                const int64_t o = w.at(i, j, k);
                const double Y = (double)(i + 1) - c1, X = (double)(j + 1) - c2, Z = (double)(k + 1) - c3;
                const double hxy = std::hypot(X, Y);  // cart2sph
                const double r = std::hypot(hxy, Z);
                const double elev = std::atan2(Z, hxy), az = std::atan2(Y, X);
                const double rc = std::cos(elev);  // sph2cart, r = 1
                const double xn = rc * std::cos(az), yn = rc * std::sin(az), zn = std::sin(elev);
                Q[(size_t)(3 * o)] = yn;
                Q[(size_t)(3 * o + 1)] = xn;
                Q[(size_t)(3 * o + 2)] = zn;
                rs[(size_t)o] = r;
                if (mask[(size_t)o] && !er[(size_t)o]) {
                    E.push_back(yn);
                    E.push_back(xn);
                    E.push_back(zn);
                    Rn.push_back(r);
                }
            }
    const int nE = (int)Rn.size();
    if (nE < 2) fail(PK_ERR_EMPTY_MASK, "OTF mask edge has fewer than two voxels");

    // knnsearch(edge, all voxels, 'K', 2) on the GPU (:135)
    DevBuf dE, dQ, dI, dD;
    alloc(dE, sizeof(double) * E.size());
    alloc(dQ, sizeof(double) * Q.size());
    alloc(dI, sizeof(int) * 2 * (size_t)nW);
    alloc(dD, sizeof(double) * 2 * (size_t)nW);
    PK_HIP(hipMemcpyAsync(dE.p, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice, stream));
    PK_HIP(hipMemcpyAsync(dQ.p, Q.data(), sizeof(double) * Q.size(), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_knn2, dim3((unsigned)((nW + 255) / 256)), dim3(256), 0, stream, dE.as<double>(), nE,
                       dQ.as<double>(), nW, dI.as<int>(), dD.as<double>());
    PK_HIP(hipGetLastError());
    std::vector<int> inds((size_t)(2 * nW));
    std::vector<double> dm((size_t)(2 * nW));
    PK_HIP(hipMemcpyAsync(inds.data(), dI.p, sizeof(int) * inds.size(), hipMemcpyDeviceToHost, stream));
    PK_HIP(hipMemcpyAsync(dm.data(), dD.p, sizeof(double) * dm.size(), hipMemcpyDeviceToHost, stream));
    PK_HIP(hipStreamSynchronize(stream));

    // rn from the two nearest edges; d_1/d_2 index d_mat rows by edge index
    // exactly as the reference does (:137-139); apodisation (:145-148)
    std::vector<double> mk((size_t)nW);
    const double kPi = 3.14159265358979323846;
    for (int64_t q = 0; q < nW; ++q) {
        const int e1 = inds[(size_t)(2 * q)], e2 = inds[(size_t)(2 * q + 1)];
        const double d1 = dm[(size_t)(2 * (int64_t)e1)], d2 = dm[(size_t)(2 * (int64_t)e2 + 1)];
        const double rn = (Rn[(size_t)e1] * d2 + Rn[(size_t)e2] * d1) / (d1 + d2);
        const double x = rs[(size_t)q] / rn;
        const double cw = std::cos(kPi * (x - lo) / 2 / (hi - lo));
        mk[(size_t)q] = (double)(x < lo) + cw * cw * (double)(x >= lo && x < hi);
    }
    std::vector<double> mfin((size_t)P, 0.0);
    std::vector<float> mcount((size_t)P, 0.f);
    const int nb = skewed ? 3 : 1;  // :150-153 sum the three z blocks
    for (int b = 0; b < nb; ++b)
        for (int64_t i = 0; i < P; ++i) {
            mfin[(size_t)i] += mk[(size_t)(i + b * P)];
            mcount[(size_t)i] += mask[(size_t)(i + b * P)] ? 1.f : 0.f;
        }
    // OTF_bp_omw = ifftshift(mask) .* conj(OTF) ./ (|OTF|.^2 + alpha) (:155-161), single
    const int64_t ish[3] = {-(pd[0] / 2), -(pd[1] / 2), -(pd[2] / 2)};
    std::vector<double> om = shift3(mfin, v, ish);
    const float al = (float)alpha;
    std::vector<cd> bp((size_t)P);
    std::vector<std::complex<float>> omw((size_t)P);
    for (int64_t i = 0; i < P; ++i) {
        const std::complex<float> o = OTF[(size_t)i];
        const float den = absO[(size_t)i] * absO[(size_t)i] + al;
        const std::complex<float> wf(o.real() / den, -o.imag() / den);
        const float mf = (float)om[(size_t)i];
        omw[(size_t)i] = std::complex<float>(mf * wf.real(), mf * wf.imag());
        bp[(size_t)i] = cd(omw[(size_t)i].real(), omw[(size_t)i].imag());
    }
    // b_omw = fftshift(real(ifftn(OTF_bp_omw))) (:163)
    fft3_host(bp, pd, true);
    std::vector<float> br((size_t)P);
    for (int64_t i = 0; i < P; ++i) br[(size_t)i] = (float)bp[(size_t)i].real();
    std::vector<float> bo = shift3(br, v, fsh);
    std::memcpy(b_out, bo.data(), sizeof(float) * (size_t)P);
    if (otf_bp_out) std::memcpy(otf_bp_out, omw.data(), sizeof(float) * 2 * (size_t)P);
    if (abs_otf_c_out) std::memcpy(abs_otf_c_out, absOc.data(), sizeof(float) * (size_t)P);
    if (mask_out) std::memcpy(mask_out, mcount.data(), sizeof(float) * (size_t)P);
    if (skewed_used) *skewed_used = skewed ? 1 : 0;
}

}  // namespace pk

// ==================================================================== Wiener-Butterworth back projector
// BackProjector(PSF_fp, bp_type, alpha, beta, n, resFlag, iRes)
// (third_parties/WBDeconvolution/BackProjector.m:1-233).  Host code: it runs
// once per PSF on PSF-sized arrays (odd and prime sizes included).  The
// reference computes the FWHMs, the cutoff frequencies and the Butterworth
// mask in single precision (its PSF is single) and the spectra in double;
// this follows the same split.
namespace pk {

// fwhm(x, y) with x = 1:N (BackProjector.m:474-509).  false: the leading
// crossing search runs off the end (the reference raises an index error);
// width NaN: no trailing crossing ("step-like pulse").
static bool wb_fwhm(const std::vector<float>& yin, float& width) {
    const int N = (int)yin.size();
    if (N < 2) return false;
    float mx = yin[0];
    for (float v : yin) mx = v > mx ? v : mx;
    std::vector<float> y((size_t)N);
    for (int i = 0; i < N; ++i) y[(size_t)i] = yin[(size_t)i] / mx;
    const float lev = 0.5f;
    auto sgn = [](float v) { return (v > 0.f) - (v < 0.f); };
    int center = 0;  // 0-based index of the first max (or min when y(1) >= 0.5)
    for (int i = 1; i < N; ++i)
        if (y[0] < lev ? y[(size_t)i] > y[(size_t)center] : y[(size_t)i] < y[(size_t)center]) center = i;
    int i = 1;
    while (sgn(y[(size_t)i] - lev) == sgn(y[(size_t)i - 1] - lev)) {
        if (++i >= N) return false;
    }
    const float ip1 = (lev - y[(size_t)i - 1]) / (y[(size_t)i] - y[(size_t)i - 1]);
    const float tlead = (float)i + ip1;  // x(i-1) + interp * (x(i) - x(i-1)), x = 1:N
    i = center + 1;
    if (i > N - 1) return false;
    while (i <= N - 2 && sgn(y[(size_t)i] - lev) == sgn(y[(size_t)i - 1] - lev)) ++i;
    if (i != N - 1) {
        const float ip2 = (lev - y[(size_t)i - 1]) / (y[(size_t)i] - y[(size_t)i - 1]);
        width = ((float)i + ip2) - tlead;
    } else {
        width = std::numeric_limits<float>::quiet_NaN();
    }
    return true;
}

// fwhm_PSF(PSF) with cFlag = 0, fitFlag = 0, pixelSize = 1 (:257-379): lines
// through the first maximum (column-major order).
static void wb_fwhm_psf(const std::vector<float>& p, const int64_t d[3], float f[3]) {
    auto need = [](bool ok) {
        if (!ok) fail(PK_ERR_INVALID_ARG, "fwhm: the PSF profile has no half-maximum crossing (index exceeds)");
    };
    f[0] = f[1] = f[2] = 0.f;
    const Vol3 v{{d[0], d[1], d[2]}};
    if (d[0] == 1 || d[1] == 1) {  // 1-D input: the whole array is the profile
        need(wb_fwhm(p, f[0]));
        return;
    }
    size_t im = 0;
    for (size_t i = 1; i < p.size(); ++i)
        if (p[i] > p[im]) im = i;
    const int64_t ix = (int64_t)im % d[0], iy = ((int64_t)im / d[0]) % d[1], iz = (int64_t)im / (d[0] * d[1]);
    std::vector<float> line;
    line.resize((size_t)d[0]);
    for (int64_t i = 0; i < d[0]; ++i) line[(size_t)i] = p[(size_t)v.at(i, iy, iz)];
    need(wb_fwhm(line, f[0]));
    line.resize((size_t)d[1]);
    for (int64_t j = 0; j < d[1]; ++j) line[(size_t)j] = p[(size_t)v.at(ix, j, iz)];
    need(wb_fwhm(line, f[1]));
    if (d[2] > 1) {
        line.resize((size_t)d[2]);
        for (int64_t k = 0; k < d[2]; ++k) line[(size_t)k] = p[(size_t)v.at(ix, iy, k)];
        need(wb_fwhm(line, f[2]));
    }
}

static double mround(double x) { return x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5); }  // MATLAB round

void wb_backprojector(const float* psf_in, const int64_t d[3], int bp_type, double alpha, double beta, int n,
                      int res_flag, const double ires[3], double* psf_bp, double* otf_bp, double* info) {
    const Vol3 v{{d[0], d[1], d[2]}};
    const int64_t P = v.n();
    if (P < 1) fail(PK_ERR_INVALID_ARG, "empty PSF");
    if (bp_type < PK_BP_TRADITIONAL || bp_type > PK_BP_WIENER_BUTTERWORTH)
        fail(PK_ERR_INVALID_ARG, "bp_type does not match any back-projector type");
    if (res_flag < 0 || res_flag > 2) fail(PK_ERR_INVALID_ARG, "Processing terminated, please set resFlag as 0, 1, or 2");
    if (d[0] == 1 || d[1] == 1) fail(PK_ERR_INVALID_ARG, "BackProjector needs a 2-D or 3-D PSF");
    std::vector<float> psf(psf_in, psf_in + P);  // single(PSF_fp)
    const double Sc[3] = {(d[0] + 1) / 2.0, (d[1] + 1) / 2.0, (d[2] + 1) / 2.0};
    const int64_t Soz = (int64_t)mround((d[2] + 1) / 2.0);
    float fw[3];
    wb_fwhm_psf(psf, d, fw);

    // traditional back projector: flipped PSF (double) and its OTF
    std::vector<cd> otf_flip((size_t)P);
    std::vector<double> flipped((size_t)P);
    for (int64_t k = 0; k < d[2]; ++k)
        for (int64_t j = 0; j < d[1]; ++j)
            for (int64_t i = 0; i < d[0]; ++i)
                flipped[(size_t)v.at(i, j, k)] = (double)psf[(size_t)v.at(d[0] - 1 - i, d[1] - 1 - j, d[2] - 1 - k)];
    const int64_t fsh[3] = {d[0] / 2, d[1] / 2, d[2] / 2}, ish[3] = {-(d[0] / 2), -(d[1] / 2), -(d[2] / 2)};
    {
        std::vector<double> s = shift3(flipped, v, ish);
        for (int64_t i = 0; i < P; ++i) otf_flip[(size_t)i] = cd(s[(size_t)i], 0.0);
        fft3_host(otf_flip, d, false);
    }
    std::vector<double> oabs((size_t)P);
    for (int64_t i = 0; i < P; ++i) oabs[(size_t)i] = std::abs(otf_flip[(size_t)i]);
    oabs = shift3(oabs, v, fsh);
    double M = oabs[0];
    for (double a : oabs) M = a > M ? a : M;
    for (auto& a : oabs) a /= M;  // OTF_abs_norm

    // resolution cutoff (single unless given), cutoff in pixels of the OTF
    float t[3];
    double td[3];
    for (int a = 0; a < 3; ++a) {
        const float px = (float)(1.0 / (double)d[a]);
        if (res_flag == 2) {
            td[a] = 1.0 / ires[a] / (1.0 / (double)d[a]);
        } else {
            const float res = res_flag == 0 ? fw[a] / (float)std::sqrt(2.0) : fw[a];
            t[a] = (1.0f / res) / px;
            td[a] = (double)t[a];
        }
    }
    // 1-based index of an OTF profile sample: max(round(Sc - t), 1) / min(round(Sc + t), S).
    // MATLAB's max and min skip NaN, so a NaN cutoff (a step-like profile has
    // no FWHM) takes index 1 / S.
    auto lo_idx = [&](int a) {
        const double x = res_flag == 2 ? mround(Sc[a] - td[a]) : mround((double)((float)Sc[a] - t[a]));
        return std::isnan(x) ? (int64_t)1 : (int64_t)std::max(x, 1.0);
    };
    auto hi_idx = [&](int a) {
        const double x = res_flag == 2 ? mround(Sc[a] + td[a]) : mround((double)((float)Sc[a] + t[a]));
        return std::isnan(x) ? d[a] : (int64_t)std::min(x, (double)d[a]);
    };
    // cutoff gains of the traditional back projector (:117-137)
    std::vector<double> lx((size_t)d[0], -1e300), ly((size_t)d[1], -1e300), lz((size_t)d[2], -1e300);
    for (int64_t k = 0; k < d[2]; ++k)
        for (int64_t j = 0; j < d[1]; ++j)
            for (int64_t i = 0; i < d[0]; ++i) {
                const double a = oabs[(size_t)v.at(i, j, k)];
                lx[(size_t)i] = std::max(lx[(size_t)i], a);
                ly[(size_t)j] = std::max(ly[(size_t)j], a);
                lz[(size_t)k] = std::max(lz[(size_t)k], a);
            }
    const double bx = (lx[(size_t)lo_idx(0) - 1] + lx[(size_t)hi_idx(0) - 1]) / 2;
    const double by = (ly[(size_t)lo_idx(1) - 1] + ly[(size_t)hi_idx(1) - 1]) / 2;
    double bz;
    if (d[2] > 1) {
        bz = (lz[(size_t)lo_idx(2) - 1] + lz[(size_t)hi_idx(2) - 1]) / 2;
    } else {
        // 2-D PSF: squeeze(max(OTF_abs_norm, [], 1)) stays a 1 x Sy row, so the
        // z "profile" is max over x of the OTF magnitude along y; both z
        // indices are 1 (:131-135)
        double m = -1e300;
        for (int64_t i = 0; i < d[0]; ++i) m = std::max(m, oabs[(size_t)v.at(i, 0, 0)]);
        bz = m;
    }
    const double beta_fp = (bx + by + bz) / 3;
    if (alpha == 1) alpha = beta_fp;
    if (beta == 1) beta = beta_fp;

    // Butterworth mask 1 / sqrt(1 + ee w^n), w = sum ((i - Sc) / t)^2, in single (:167-177, :213-222)
    auto butterworth = [&](double ee) {
        std::vector<double> m((size_t)P);
        const float eef = (float)ee, nf = (float)n;
        for (int64_t k = 0; k < d[2]; ++k)
            for (int64_t j = 0; j < d[1]; ++j)
                for (int64_t i = 0; i < d[0]; ++i) {
                    float w;
                    if (res_flag == 2) {
                        const double a = (i + 1 - Sc[0]) / td[0], b = (j + 1 - Sc[1]) / td[1], c = (k + 1 - Sc[2]) / td[2];
                        w = (float)(a * a + b * b + c * c);
                    } else {
                        const float a = (float)(i + 1 - Sc[0]) / t[0], b = (float)(j + 1 - Sc[1]) / t[1],
                                    c = (float)(k + 1 - Sc[2]) / t[2];
                        w = a * a + b * b + c * c;
                    }
                    m[(size_t)v.at(i, j, k)] = (double)(1.0f / std::sqrt(1.0f + eef * std::pow(w, nf)));
                }
        return shift3(m, v, ish);  // ifftshift
    };
    std::vector<cd> otf((size_t)P);
    std::vector<double> out((size_t)P);
    auto psf_from_otf = [&]() {  // PSF_bp = fftshift(real(ifftn(OTF_bp)))
        std::vector<cd> tmp = otf;
        fft3_host(tmp, d, true);
        std::vector<double> r((size_t)P);
        for (int64_t i = 0; i < P; ++i) r[(size_t)i] = tmp[(size_t)i].real();
        out = shift3(r, v, fsh);
    };
    double beta_w = 0;
    switch (bp_type) {
        case PK_BP_TRADITIONAL:
            out = flipped;
            otf = otf_flip;
            break;
        case PK_BP_GAUSSIAN: {  // gen_gaussianPSF_3D(Sx, Sy, Sz, FWHMx, FWHMy, FWHMz), single (:511-533)
            const float sx = fw[0] / 2.3548f, sy = fw[1] / 2.3548f, sz = fw[2] / 2.3548f;
            const float qx = sx * sx * 2, qy = sy * sy * 2, qz = sz * sz * 2;
            const float coef = 1.0f / ((float)std::pow(2 * kPi, 1.5) * sx * sy * sz);
            for (int64_t k = 0; k < d[2]; ++k)
                for (int64_t j = 0; j < d[1]; ++j)
                    for (int64_t i = 0; i < d[0]; ++i) {
                        const float a = (float)(i + 1 - Sc[0]), b = (float)(j + 1 - Sc[1]), c = (float)(k + 1 - Sc[2]);
                        const float dd = a * a / qx + b * b / qy + c * c / qz;
                        out[(size_t)v.at(i, j, k)] = (double)(coef * std::exp(-dd));
                    }
            std::vector<double> s = shift3(out, v, ish);
            for (int64_t i = 0; i < P; ++i) otf[(size_t)i] = cd(s[(size_t)i], 0.0);
            fft3_host(otf, d, false);
            break;
        }
        case PK_BP_BUTTERWORTH: {
            std::vector<double> m = butterworth(1.0 / (beta * beta) - 1.0);
            for (int64_t i = 0; i < P; ++i) otf[(size_t)i] = cd(m[(size_t)i], 0.0);
            psf_from_otf();
            break;
        }
        case PK_BP_WIENER:
        case PK_BP_WIENER_BUTTERWORTH: {
            std::vector<cd> wn((size_t)P);
            for (int64_t i = 0; i < P; ++i) {
                const cd o = otf_flip[(size_t)i] / M;  // OTF_flip_norm
                const double a = std::abs(o);
                wn[(size_t)i] = o / (a * a + alpha);
            }
            if (bp_type == PK_BP_WIENER) {
                otf = wn;
            } else {
                // cutoff gain of the Wiener filter on the central z slice (:207-212)
                std::vector<double> wa((size_t)P);
                for (int64_t i = 0; i < P; ++i) wa[(size_t)i] = std::abs(wn[(size_t)i]);
                wa = shift3(wa, v, fsh);
                std::vector<double> tl((size_t)d[0], -1e300);
                for (int64_t j = 0; j < d[1]; ++j)
                    for (int64_t i = 0; i < d[0]; ++i)
                        tl[(size_t)i] = std::max(tl[(size_t)i], wa[(size_t)v.at(i, j, Soz - 1)]);
                beta_w = (tl[(size_t)lo_idx(0) - 1] + tl[(size_t)hi_idx(0) - 1]) / 2;
                std::vector<double> m = butterworth(beta_w / (beta * beta) - 1.0);
                for (int64_t i = 0; i < P; ++i) otf[(size_t)i] = m[(size_t)i] * wn[(size_t)i];
            }
            psf_from_otf();
            break;
        }
    }
    std::memcpy(psf_bp, out.data(), sizeof(double) * (size_t)P);
    if (otf_bp) std::memcpy(otf_bp, otf.data(), sizeof(double) * 2 * (size_t)P);
    if (info) {
        const double vals[10] = {fw[0], fw[1], fw[2], td[0], td[1], td[2], beta_fp, alpha, beta, beta_w};
        std::memcpy(info, vals, sizeof(vals));
    }
}

}  // namespace pk

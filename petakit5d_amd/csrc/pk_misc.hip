// pk_misc.hip -- elementwise and bookkeeping kernels of the deconvolution
// engine (everything that is not a transform pass).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "pk_engine.hpp"
#include "pk_misc.hpp"

namespace pk {

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Zero-padded, circularly shifted PSF (decon_psf2otf.m:19-24): psf(i,j,k)
// lands at ((i - floor(p1/2)) mod n1, (j - floor(p2/2)) mod n2, (k - floor(p3/2)) mod n3).
__global__ void k_scatter_psf(const float* __restrict__ q, int p1, int p2, int p3, float* __restrict__ dst,
                              int64_t n1, int64_t n2, int64_t n3) {
    const int64_t P = (int64_t)p1 * p2 * p3;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < P; idx += stride) {
        const int i = (int)(idx % p1);
        const int j = (int)((idx / p1) % p2);
        const int k = (int)(idx / ((int64_t)p1 * p2));
        int64_t x = (i - p1 / 2) % n1;
        if (x < 0) x += n1;
        int64_t y = (j - p2 / 2) % n2;
        if (y < 0) y += n2;
        int64_t z = (k - p3 / 2) % n3;
        if (z < 0) z += n3;
        dst[x + n1 * (y + n2 * z)] = q[idx];
    }
}

// Biggs step length from the block partials of the previous update, summed
// in a fixed order (one block): lambda = num / (den + eps), clamped to
// [0, 1] with MATLAB min/max NaN rules (decon_lucy_function.m:126-127).
__global__ void k_lambda(const double* __restrict__ partials, int nblk, float eps, float* __restrict__ lambda) {
    __shared__ double sn[256], sd[256];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int i = t; i < nblk; i += 256) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    sn[t] = a;
    sd[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sn[t] += sn[t + w];
            sd[t] += sd[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const float num = (float)sn[0];
        const float den = (float)sd[0];
        float l = num / (den + eps);
        l = (l > 1.f || l != l) ? 1.f : l;  // min(lambda, 1): min(NaN, 1) = 1
        l = l > 0.f ? l : 0.f;               // max(lambda, 0)
        *lambda = l;
    }
}

// Epilogue (decon_lucy_function.m:177-203): damp, scale / offset / mask,
// uint16 conversion and the 1-based inclusive bbox crop, one pass over the
// output voxels.  Work items are (output row, 1024-voxel chunk) pairs: the
// row decomposition is block-uniform, so no per-voxel integer division; when
// the crop keeps whole dim-1 x dim-2 planes the output is one contiguous run.
__global__ void __launch_bounds__(256) k_epilogue(EpiArgs a) {
    const bool flat = (a.o1 == a.n1 && a.o2 == a.n2);
    const int64_t rlen = flat ? a.o1 * a.o2 * a.o3 : a.o1;
    const int64_t nrows = flat ? 1 : a.o2 * a.o3;
    const int64_t chunks = (rlen + kEpiChunk - 1) / kEpiChunk;
    const int64_t total = nrows * chunks;
    // whole planes kept and 4-voxel aligned: 16-byte loads, 8 / 16-byte stores
    const uintptr_t al = reinterpret_cast<uintptr_t>(a.J) | reinterpret_cast<uintptr_t>(a.I) |
                         (a.out16 ? (reinterpret_cast<uintptr_t>(a.out16) & 7u) : reinterpret_cast<uintptr_t>(a.outf));
    const bool vec = flat && (rlen % 4 == 0) && ((a.b3 * a.n1 * a.n2) % 4 == 0) && (al % 16 == 0);
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int64_t r = w / chunks;
        const int64_t x0 = (w - r * chunks) * kEpiChunk;
        int64_t src, dst;
        if (flat) {
            src = a.b3 * a.n1 * a.n2;
            dst = 0;
        } else {
            const int64_t y = r % a.o2, z = r / a.o2;
            src = a.b1 + a.n1 * ((y + a.b2) + a.n2 * (z + a.b3));
            dst = r * a.o1;
        }
        if (vec) {
            const int64_t x = x0 + 4 * (int64_t)threadIdx.x;
            if (x >= rlen) continue;
            const int64_t i = src + x;
            const float4 jv = *reinterpret_cast<const float4*>(a.J + i);
            float4 dv = make_float4(0.f, 0.f, 0.f, 0.f);
            if (a.damp > 1.f) dv = *reinterpret_cast<const float4*>(a.I + i);
            const int64_t mi = a.mask_at_out ? dst + x : i;  // mask index
            const float o0 = epi_voxel(a, jv.x, dv.x, epi_mask(a, mi));
            const float o1 = epi_voxel(a, jv.y, dv.y, epi_mask(a, mi + 1));
            const float o2 = epi_voxel(a, jv.z, dv.z, epi_mask(a, mi + 2));
            const float o3 = epi_voxel(a, jv.w, dv.w, epi_mask(a, mi + 3));
            if (a.out16) {
                ushort4 u;
                u.x = (uint16_t)u16_round(o0);
                u.y = (uint16_t)u16_round(o1);
                u.z = (uint16_t)u16_round(o2);
                u.w = (uint16_t)u16_round(o3);
                *reinterpret_cast<ushort4*>(a.out16 + dst + x) = u;
            } else {
                *reinterpret_cast<float4*>(a.outf + dst + x) = make_float4(o0, o1, o2, o3);
            }
            continue;
        }
        float jv[kEpiChunk / 256], dv[kEpiChunk / 256], mv[kEpiChunk / 256];
#pragma unroll
        for (int k = 0; k < kEpiChunk / 256; ++k) {
            const int64_t x = x0 + k * 256 + threadIdx.x;
            const bool in = x < rlen;
            const int64_t i = src + (in ? x : 0);
            jv[k] = a.J[i];
            dv[k] = a.damp > 1.f ? a.I[i] : 0.f;
            mv[k] = epi_mask(a, a.mask_at_out ? dst + (in ? x : 0) : i);
        }
#pragma unroll
        for (int k = 0; k < kEpiChunk / 256; ++k) {
            const int64_t x = x0 + k * 256 + threadIdx.x;
            if (x >= rlen) continue;
            const float j = epi_voxel(a, jv[k], dv[k], mv[k]);
            if (a.out16) a.out16[dst + x] = (uint16_t)u16_round(j);
            else a.outf[dst + x] = j;
        }
    }
}

// ---- debug metric (decon_lucy_function.m:145): sum((double(J) - double(I)).^2)
// as gridDim.x block partials (grid-stride, then a fixed-order tree), reduced
// by one block in a fixed order: deterministic.
__global__ void k_sqdiff_partial(const float* __restrict__ J, const float* __restrict__ I, int64_t n,
                                 double* __restrict__ partials) {
    __shared__ double sh[256];
    double a = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double d = (double)J[i] - (double)I[i];
        a += d * d;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void k_sum_partials(const double* __restrict__ partials, int n, double* __restrict__ out) {
    __shared__ double sh[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += partials[i];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ---- edge-erosion mask (decon_mask_edge_erosion.m)
__global__ void k_mask_init(const uint16_t* __restrict__ raw16, const float* __restrict__ rawf, int64_t n,
                            uint8_t* __restrict__ m, unsigned long long* __restrict__ count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint8_t v = raw16 ? (raw16[i] != 0) : (rawf[i] != 0.f);
        m[i] = v;
        c += v;
    }
    // integer counts: order-independent, so the atomic keeps results deterministic
    atomicAdd(count, c);
}

// Full-rectangle mask: clear an E-voxel border on every face (:8-15); else
// clear the outer 1-voxel faces (:18-21).
__global__ void k_mask_faces(uint8_t* __restrict__ m, int64_t n1, int64_t n2, int64_t n3, int is3d,
                             const unsigned long long* __restrict__ count, int E) {
    const int64_t n = n1 * n2 * n3;
    const bool full = (*count == (unsigned long long)n);
    const int w = full ? E : 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t y = i % n1, x = (i / n1) % n2, z = i / (n1 * n2);
        bool clr = (y < w || y >= n1 - w || x < w || x >= n2 - w);
        if (is3d && (z < w || z >= n3 - w)) clr = true;
        if (clr) m[i] = 0;
    }
}

// Neighbourhood of strel('disk', r) with MATLAB's default N = 4, the element
// decon_mask_edge_erosion.m:22-27 erodes with.  r < 3: the exact disk
// dx^2 + dy^2 <= r^2.  r >= 3: the radial decomposition into four periodic
// lines documented for strel (Adams 1993; Jones & Soille 1996): offsets
// (1,0), (1,1), (0,1), (-1,1), each repeated floor(k / |v|) times either side,
// k = 2r / (cot(pi/8) + 1 / sin(pi/8)), then -- the sum falling short of r --
// a horizontal and a vertical line of length 2(r - e - 1) + 1 when that is
// >= 3 (e: the sum's row extent).  The element is convex and symmetric, so it
// is kept as the half-width of each row.
DiskSE disk_se(int r) {
    DiskSE se{};
    if (r < 0 || r > kDiskMaxR) fail(PK_ERR_INVALID_ARG, "EdgeErosion - 1 must lie in [0, " +
                                                             std::to_string(kDiskMaxR) + "]");
    if (r < 3) {
        se.R = r;
        for (int dy = -r; dy <= r; ++dy) {
            int h = 0;
            while ((h + 1) * (h + 1) + dy * dy <= r * r) ++h;
            se.w[dy + r] = (int16_t)h;
        }
        return se;
    }
    static const int v[4][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}};
    const double t = 3.14159265358979323846 / 8.0;
    const double k = 2.0 * r / (1.0 / std::tan(t) + 1.0 / std::sin(t));
    // Minkowski sum of the lines, row by row: every row of every partial sum
    // is one run [lo, hi] (a line's copies of a run overlap or touch), on a
    // (2r+1)-row table (the sum never reaches past r)
    const int G = 2 * r + 1;
    const int kNone = 1 << 30;
    std::vector<int> lo(G, kNone), hi(G, -kNone), nlo(G), nhi(G);
    lo[r] = hi[r] = 0;
    for (int q = 0; q < 4; ++q) {
        const int rp = (int)std::floor(k / std::sqrt((double)(v[q][0] * v[q][0] + v[q][1] * v[q][1])));
        std::fill(nlo.begin(), nlo.end(), kNone);
        std::fill(nhi.begin(), nhi.end(), -kNone);
        for (int y = 0; y < G; ++y) {
            if (lo[y] > hi[y]) continue;
            for (int i = -rp; i <= rp; ++i) {
                const int yy = y + i * v[q][0];
                if (yy < 0 || yy >= G) continue;
                nlo[yy] = std::min(nlo[yy], lo[y] + i * v[q][1]);
                nhi[yy] = std::max(nhi[yy], hi[y] + i * v[q][1]);
            }
        }
        lo.swap(nlo);
        hi.swap(nhi);
    }
    std::vector<int> w(G, -1);  // the full sum is symmetric in dx: lo = -hi
    for (int y = 0; y < G; ++y)
        if (lo[y] <= hi[y]) w[y] = hi[y];
    int e = 0;
    for (int y = 0; y < G; ++y)
        if (w[y] >= 0) e = std::max(e, std::abs(y - r));
    const int len = 2 * (r - e - 1) + 1;
    const int h = len >= 3 ? (len - 1) / 2 : 0;
    se.R = e + h;
    for (int dy = -se.R; dy <= se.R; ++dy) {
        int best = -1;
        for (int sy = dy - h; sy <= dy + h; ++sy)
            if (sy >= -e && sy <= e) best = std::max(best, w[sy + r]);
        se.w[dy + se.R] = (int16_t)(best + h);
    }
    return se;
}

// In-plane erosion of every slice by the element se (row dy in [-R, R] spans
// |dx| <= w[dy + R]); out-of-image neighbours count as set (binary erosion
// padding, once around the whole decomposed element).
__global__ void k_mask_erode_disk(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1,
                                  int64_t n2, int64_t n3, DiskSE se, const unsigned long long* __restrict__ count) {
    const int64_t n = n1 * n2 * n3;
    const bool full = (*count == (unsigned long long)n);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (full) {
            dst[i] = src[i];
            continue;
        }
        const int64_t y = i % n1, x = (i / n1) % n2, z = i / (n1 * n2);
        uint8_t v = 1;
        for (int dy = -se.R; dy <= se.R && v; ++dy) {
            const int64_t yy = y + dy;
            if (yy < 0 || yy >= n1) continue;
            const int hw = se.w[dy + se.R];
            const int64_t x0 = x - hw < 0 ? 0 : x - hw, x1 = x + hw >= n2 ? n2 - 1 : x + hw;
            for (int64_t xx = x0; xx <= x1 && v; ++xx) v = src[yy + n1 * (xx + n2 * z)];
        }
        dst[i] = v;
    }
}

// ---- whole-volume edge erosion of the large-volume driver
// (RLdecon_large_in_memory.m:150-155): bw = im > 0 with the six faces cleared,
// kept where bwdist(~bw) > E - 1, i.e. where no false voxel lies within
// Euclidean distance E - 1.  The squared distance to the nearest false voxel
// separates over the axes (d^2 = min over false voxels of dy^2 + dx^2 + dz^2),
// and only distances up to R = E - 1 matter, so three passes of a bounded
// search give it exactly: pass 1 along dim 1 from the mask, passes 2 and 3
// along dims 2 and 3 over the previous pass's squared distances, each taking
// min(d_prev(q + k) + k^2) over |k| <= R; values above R^2 saturate at 65535.
__global__ void k_edt_mask_pass1(const uint16_t* __restrict__ raw16, const float* __restrict__ rawf, int64_t n1,
                                 int64_t n2, int64_t n3, int R, uint16_t* __restrict__ d) {
    const int64_t n = n1 * n2 * n3;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int R2 = R * R;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t y = i % n1, x = (i / n1) % n2, z = i / (n1 * n2);
        const int64_t row = i - y;
        auto bw = [&](int64_t yy) -> bool {  // im > 0, faces cleared
            if (yy == 0 || yy == n1 - 1 || x == 0 || x == n2 - 1 || z == 0 || z == n3 - 1) return false;
            return raw16 ? raw16[row + yy] > 0 : rawf[row + yy] > 0.f;
        };
        int best = 65535;
        for (int k = 0; k <= R && best == 65535; ++k) {  // nearest false voxel along the row
            if ((y - k >= 0 && !bw(y - k)) || (y + k < n1 && !bw(y + k))) best = k * k;
        }
        d[i] = (uint16_t)(best <= R2 ? best : 65535);
    }
}

__global__ void k_edt_mask_pass(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst, int64_t n1, int64_t n2,
                                int64_t n3, int axis, int R) {
    const int64_t n = n1 * n2 * n3;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int R2 = R * R;
    const int64_t step = axis == 2 ? n1 : n1 * n2;
    const int64_t len = axis == 2 ? n2 : n3;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t q = axis == 2 ? (i / n1) % n2 : i / (n1 * n2);
        int best = 65535;
        for (int k = -R; k <= R; ++k) {
            const int64_t qq = q + k;
            if (qq < 0 || qq >= len) continue;
            const int v = src[i + k * step];
            if (v == 65535) continue;
            const int c = v + k * k;
            if (c < best) best = c;
        }
        dst[i] = (uint16_t)(best <= R2 ? best : 65535);
    }
}

__global__ void k_edt_mask_final(const uint16_t* __restrict__ d, uint8_t* __restrict__ mask, int64_t n, int R) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int R2 = R * R;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        mask[i] = (uint8_t)(d[i] == 65535 || (int)d[i] > R2);
}

// Erosion along dim 3 with a (2h+1)-long line (decon_mask_edge_erosion.m:25-27).
__global__ void k_mask_erode_z(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1, int64_t n2,
                               int64_t n3, int h, const unsigned long long* __restrict__ count) {
    const int64_t n = n1 * n2 * n3;
    const bool full = (*count == (unsigned long long)n);
    const int64_t plane = n1 * n2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (full) {
            dst[i] = src[i];
            continue;
        }
        const int64_t z = i / plane, q = i - z * plane;
        uint8_t v = 1;
        for (int dz = -h; dz <= h && v; ++dz) {
            const int64_t zz = z + dz;
            if (zz < 0 || zz >= n3) continue;
            v = src[q + plane * zz];
        }
        dst[i] = v;
    }
}

}  // namespace pk

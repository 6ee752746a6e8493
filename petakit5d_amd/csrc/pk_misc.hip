// pk_misc.hip -- elementwise and bookkeeping kernels of the deconvolution
// engine (everything that is not a transform pass).
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// In-plane erosion with a disk of radius r (pixels with dy^2+dx^2 <= r^2);
// out-of-image neighbours count as set (binary erosion padding).
__global__ void k_mask_erode_disk(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1,
                                  int64_t n2, int64_t n3, int r, const unsigned long long* __restrict__ count) {
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
        for (int dx = -r; dx <= r && v; ++dx)
            for (int dy = -r; dy <= r && v; ++dy) {
                if (dx * dx + dy * dy > r * r) continue;
                const int64_t yy = y + dy, xx = x + dx;
                if (yy < 0 || yy >= n1 || xx < 0 || xx >= n2) continue;
                v = src[yy + n1 * (xx + n2 * z)];
            }
        dst[i] = v;
    }
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

// pk_misc.hpp -- declarations of the non-transform kernels.
#pragma once
#include <algorithm>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pk {

struct EpiArgs {
    const float* J;
    const float* I;
    const uint8_t* mask;      // eroded mask (EdgeErosion > 0) or null
    const uint16_t* raw16;    // raw input for I ~= 0 when no eroded mask
    const float* rawf;
    uint16_t* out16;
    float* outf;
    int64_t n1, n2;           // source volume
    int64_t o1, o2, o3;       // output (bbox) size
    int64_t b1, b2, b3;       // 0-based bbox start
    float damp, scale, offset;
    int apply_scale, use_mask;
    int pair;                 // fused form: dim-1 crop start and size even, output aligned for pair stores
    int mask_at_out;          // raw16 / rawf in the output's (unpadded) layout: the I ~= 0 mask reads there
};

// ---- per-voxel epilogue step (decon_lucy_function.m:177-203), shared by
// k_epilogue and the fused last update of the row pass, so both give the same
// bits.  Products and sums round separately, as MATLAB evaluates them.
#pragma clang fp contract(off)

__device__ __forceinline__ float u16_round(float x) {
    // MATLAB uint16(): round half away from zero, saturate, NaN -> 0
    if (!(x > 0.f)) return 0.f;
    if (x >= 65535.f) return 65535.f;
    return roundf(x);
}
__device__ __forceinline__ float epi_voxel(const EpiArgs& a, float j, float dI_raw, float m) {
    if (a.damp > 1.f) {
        const float dI = a.damp * dI_raw;
        // J.*(J<=dI) + min(J, dI.*(J>=dI)); exact equality yields 2J as in the reference
        const float t1 = (j <= dI) ? j : 0.f;
        const float t2 = (j >= dI) ? dI : 0.f;
        j = t1 + (j < t2 ? j : t2);
    }
    if (a.apply_scale) {
        if (a.use_mask) j = (j * a.scale + a.offset) * m;
        else j = j * a.scale;
    }
    return j;
}
__device__ __forceinline__ float epi_mask(const EpiArgs& a, int64_t i) {
    if (!(a.apply_scale && a.use_mask)) return 1.f;
    if (a.mask) return a.mask[i] ? 1.f : 0.f;
    if (a.raw16) return a.raw16[i] != 0 ? 1.f : 0.f;
    return a.rawf[i] != 0.f ? 1.f : 0.f;
}

// Fused form: one dim-1 row of the estimate, row index `row` over dims 2 x 3.
struct EpiRow {
    int64_t src;  // volume index of x = 0
    int64_t dst;  // output index of x = 0 (x outside the crop never stores)
    bool keep;    // the row lies inside the crop in dims 2 and 3
};
__device__ __forceinline__ EpiRow epi_row(const EpiArgs& a, int64_t row) {
    const int64_t y = row % a.n2, z = row / a.n2;
    EpiRow r;
    r.src = row * a.n1;
    r.keep = y >= a.b2 && y < a.b2 + a.o2 && z >= a.b3 && z < a.b3 + a.o3;
    r.dst = ((y - a.b2) + a.o2 * (z - a.b3)) * a.o1 - a.b1;
    return r;
}
// Values o.x (and o.y when pair) of the estimate at x0 (and x0 + 1) of row r.
__device__ __forceinline__ void epi_put(const EpiArgs& a, const EpiRow& r, int64_t x0, bool pair, float2 o) {
    float dv0 = 0.f, dv1 = 0.f;
    if (a.damp > 1.f) {
        if (pair) {
            const float2 d = *reinterpret_cast<const float2*>(a.I + r.src + x0);
            dv0 = d.x;
            dv1 = d.y;
        } else {
            dv0 = a.I[r.src + x0];
        }
    }
    const int64_t mi = (a.mask_at_out ? r.dst : r.src) + x0;  // mask index of x0
    const float v0 = epi_voxel(a, o.x, dv0, epi_mask(a, mi));
    if (!pair) {
        if (x0 < a.b1 || x0 >= a.b1 + a.o1) return;
        if (a.out16) a.out16[r.dst + x0] = (uint16_t)u16_round(v0);
        else a.outf[r.dst + x0] = v0;
        return;
    }
    const float v1 = epi_voxel(a, o.y, dv1, epi_mask(a, mi + 1));
    if (a.pair) {  // x0 even, crop start and size even: both values in or both out
        if (x0 < a.b1 || x0 >= a.b1 + a.o1) return;
        if (a.out16) {
            const uint32_t w = (uint32_t)u16_round(v0) | ((uint32_t)u16_round(v1) << 16);
            *reinterpret_cast<uint32_t*>(a.out16 + r.dst + x0) = w;
        } else {
            *reinterpret_cast<float2*>(a.outf + r.dst + x0) = make_float2(v0, v1);
        }
        return;
    }
    if (x0 >= a.b1 && x0 < a.b1 + a.o1) {
        if (a.out16) a.out16[r.dst + x0] = (uint16_t)u16_round(v0);
        else a.outf[r.dst + x0] = v0;
    }
    if (x0 + 1 >= a.b1 && x0 + 1 < a.b1 + a.o1) {
        if (a.out16) a.out16[r.dst + x0 + 1] = (uint16_t)u16_round(v1);
        else a.outf[r.dst + x0 + 1] = v1;
    }
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v);
__global__ void k_scatter_psf(const float* __restrict__ q, int p1, int p2, int p3, float* __restrict__ dst,
                              int64_t n1, int64_t n2, int64_t n3);
__global__ void k_lambda(const double* __restrict__ partials, int nblk, float eps, float* __restrict__ lambda);
// debug metric: block partials of sum((J - I)^2) in double (blockDim 256), then their fixed-order sum
__global__ void k_sqdiff_partial(const float* __restrict__ J, const float* __restrict__ I, int64_t n,
                                 double* __restrict__ partials);
__global__ void k_sum_partials(const double* __restrict__ partials, int n, double* __restrict__ out);
constexpr int kEpiChunk = 1024;  // output voxels per epilogue work item (4 per thread)
__global__ void k_epilogue(EpiArgs a);
// grid for k_epilogue: one block per work item, capped (the kernel strides)
inline unsigned epi_grid(const EpiArgs& a) {
    const bool flat = (a.o1 == a.n1 && a.o2 == a.n2);
    const int64_t rlen = flat ? a.o1 * a.o2 * a.o3 : a.o1;
    const int64_t nrows = flat ? 1 : a.o2 * a.o3;
    const int64_t total = nrows * ((rlen + kEpiChunk - 1) / kEpiChunk);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(total, 65536));
}
__global__ void k_mask_init(const uint16_t* __restrict__ raw16, const float* __restrict__ rawf, int64_t n,
                            uint8_t* __restrict__ m, unsigned long long* __restrict__ count);
__global__ void k_mask_faces(uint8_t* __restrict__ m, int64_t n1, int64_t n2, int64_t n3, int is3d,
                             const unsigned long long* __restrict__ count, int E);
// strel('disk', r) as the half-width of each row dy in [-R, R]: |dx| <= w[dy + R]
constexpr int kDiskMaxR = 255;
struct DiskSE {
    int R;
    int16_t w[2 * kDiskMaxR + 1];
};
DiskSE disk_se(int r);  // host: the element decon_mask_edge_erosion.m:22-27 uses (default N = 4)
__global__ void k_mask_erode_disk(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1,
                                  int64_t n2, int64_t n3, DiskSE se, const unsigned long long* __restrict__ count);
// bwdist(~bw) > R of bw = im > 0 with cleared faces (RLdecon_large_in_memory.m:150-155), R <= 255
__global__ void k_edt_mask_pass1(const uint16_t* __restrict__ raw16, const float* __restrict__ rawf, int64_t n1,
                                 int64_t n2, int64_t n3, int R, uint16_t* __restrict__ d);
__global__ void k_edt_mask_pass(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst, int64_t n1, int64_t n2,
                                int64_t n3, int axis, int R);
__global__ void k_edt_mask_final(const uint16_t* __restrict__ d, uint8_t* __restrict__ mask, int64_t n, int R);
__global__ void k_mask_erode_z(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1, int64_t n2,
                               int64_t n3, int h, const unsigned long long* __restrict__ count);

}  // namespace pk

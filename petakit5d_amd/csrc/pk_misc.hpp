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
};

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v);
__global__ void k_scatter_psf(const float* __restrict__ q, int p1, int p2, int p3, float* __restrict__ dst,
                              int64_t n1, int64_t n2, int64_t n3);
__global__ void k_lambda(const double* __restrict__ partials, int nblk, float eps, float* __restrict__ lambda);
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
__global__ void k_mask_erode_disk(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1,
                                  int64_t n2, int64_t n3, int r, const unsigned long long* __restrict__ count);
__global__ void k_mask_erode_z(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n1, int64_t n2,
                               int64_t n3, int h, const unsigned long long* __restrict__ count);

}  // namespace pk

// pk_deskew.hip -- deskew of a deconvolved stage-scan volume
// (pointDetection/llsmtools/transforms/deskewFrame3D.m:26-70, called from
// microscopeDataProcessing/decon/RLdecon.m:448-449).
//
// The reference shears the volume with imwarp: output column c (1-based) of
// plane p takes the input at column x = c - xstep * p - xshift, same row, same
// plane, linearly interpolated, where dx = cos(angle) dz / xyPixelSize,
// (xstep, xshift) = (dx, -dx) or (-dx, dx + ceil((nz - 1) dx)) when reversed,
// the output is ceil((nz - 1) dx) + nx columns wide and 'Crop' trims
// floor(nx / 2) (when that width reaches 2 nx) or floor(ceil((nz - 1) dx) / 2)
// columns off each side.  imwarp pads the input with one column of the fill
// value (0) on each side, so positions in [0, 1) and (nx, nx + 1] blend the edge
// column with 0 and positions further out are 0.
//
// MATLAB volumes are column-major: a column (fixed x, z) is ny contiguous
// floats, and every value of an output column uses the same two input columns
// and weight.  One block per output column computes the source position once
// and walks down the column: coalesced reads and writes, HBM-bound (4 bytes
// read + 4 written per value; the second input column comes from cache).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pk_engine.hpp"

namespace pk {

struct Affine3 {
    double a[4][4];  // row-vector convention: [x y z 1] = [u v w 1] * a
};

struct DeskewGeom {
    int64_t nx_out;
    double xstep, xshift;
};

DeskewGeom deskew_geometry(const int64_t dims[3], double angle, double dz, double xy, bool reverse, bool crop) {
    const int64_t nx = dims[1], nz = dims[2];
    const double dx = std::cos(angle * M_PI / 180.0) * dz / xy;
    const double span = std::ceil((double)(nz - 1) * dx);
    DeskewGeom g;
    g.xstep = reverse ? -dx : dx;
    g.xshift = reverse ? dx + span : -dx;
    g.nx_out = (int64_t)span + nx;
    if (crop) {
        const int64_t cw = g.nx_out >= 2 * nx ? nx / 2 : (int64_t)std::floor(span / 2.0);
        g.xshift -= (double)cw;
        g.nx_out -= 2 * cw;
    }
    return g;
}

// VEC: ny % 4 == 0 and 16-byte aligned buffers, so every column is whole
// float4s (one 16-byte load per operand and one store per four values).
template <bool VEC>
__global__ void __launch_bounds__(256) k_deskew(const float* __restrict__ vol, float* __restrict__ out, int64_t ny,
                                                int64_t nx, int64_t nx_out, double xstep, double xshift) {
    // one block per output column (c, p); its threads walk down the column
    const int64_t col = blockIdx.x;
    const int64_t c = col % nx_out, p = col / nx_out;
    const double x = (double)(c + 1) - xstep * (double)(p + 1) - xshift;  // 1-based input column
    float* __restrict__ o = out + col * ny;
    if (!(x >= 0.0 && x <= (double)(nx + 1))) {
        if constexpr (VEC) {
            for (int64_t r = threadIdx.x; r < ny / 4; r += blockDim.x)
                reinterpret_cast<float4*>(o)[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            for (int64_t r = threadIdx.x; r < ny; r += blockDim.x) o[r] = 0.f;
        }
        return;
    }
    const double x0 = std::floor(x);
    const float f = (float)(x - x0);
    const int64_t i0 = (int64_t)x0;  // 0 ... nx + 1
    const float* __restrict__ plane = vol + p * ny * nx;
    const float* __restrict__ ca = (i0 >= 1 && i0 <= nx) ? plane + (i0 - 1) * ny : nullptr;
    const float* __restrict__ cb = (i0 + 1 >= 1 && i0 + 1 <= nx) ? plane + i0 * ny : nullptr;
    if constexpr (VEC) {
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int64_t r = threadIdx.x; r < ny / 4; r += blockDim.x) {
            const float4 a = ca ? reinterpret_cast<const float4*>(ca)[r] : z;
            const float4 b = cb ? reinterpret_cast<const float4*>(cb)[r] : z;
            reinterpret_cast<float4*>(o)[r] =
                make_float4(a.x + f * (b.x - a.x), a.y + f * (b.y - a.y), a.z + f * (b.z - a.z), a.w + f * (b.w - a.w));
        }
    } else {
        for (int64_t r = threadIdx.x; r < ny; r += blockDim.x) {
            const float a = ca ? ca[r] : 0.f;
            const float b = cb ? cb[r] : 0.f;
            o[r] = a + f * (b - a);
        }
    }
}

void launch_deskew(const float* vol, float* out, const int64_t dims[3], const DeskewGeom& g, hipStream_t s) {
    const int64_t cols = g.nx_out * dims[2];
    if (cols <= 0 || dims[0] <= 0) return;
    const bool vec = dims[0] % 4 == 0 && reinterpret_cast<uintptr_t>(vol) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(out) % 16 == 0;
    const int64_t per_col = vec ? dims[0] / 4 : dims[0];  // threads' work items per column
    const unsigned nt = per_col >= 256 ? 256u : (unsigned)((per_col + 63) / 64 * 64);
    // the dispatch's grid size counts work-items in 32 bits
    if (cols * (int64_t)nt > 0xFFFFFFFFll) fail(PK_ERR_UNSUPPORTED_SIZE, "deskewed volume has too many columns");
    if (vec)
        hipLaunchKernelGGL(k_deskew<true>, dim3((unsigned)cols), dim3(nt), 0, s, vol, out, dims[0], dims[1], g.nx_out,
                           g.xstep, g.xshift);
    else
        hipLaunchKernelGGL(k_deskew<false>, dim3((unsigned)cols), dim3(nt), 0, s, vol, out, dims[0], dims[1],
                           g.nx_out, g.xstep, g.xshift);
    PK_HIP(hipGetLastError());
}

}  // namespace pk

namespace pk {

// Affine resampling of a column-major volume, imwarp-style: output voxel
// (row r, column c, page p), 1-based world point (u, v, w) = (c, r, p), reads
// the input at [x y z 1] = [u v w 1] * inv_m (row-vector convention, x the
// column, y the row, z the page), trilinear over the input padded with one
// voxel of the fill value 0 on every side (rotateFrame3D.m:46-100 resamples
// through imwarp this way).  One block per output column; threads walk down it.
__global__ void __launch_bounds__(256) k_affine3d(const float* __restrict__ vol, float* __restrict__ out, int64_t ny,
                                                  int64_t nx, int64_t nz, int64_t oy, int64_t ox, Affine3 m) {
    const int64_t col = blockIdx.x;
    const int64_t c = col % ox, p = col / ox;
    const double u = (double)(c + 1), w = (double)(p + 1);
    const double bx = u * m.a[0][0] + w * m.a[2][0] + m.a[3][0];
    const double by = u * m.a[0][1] + w * m.a[2][1] + m.a[3][1];
    const double bz = u * m.a[0][2] + w * m.a[2][2] + m.a[3][2];
    float* __restrict__ o = out + col * oy;
    for (int64_t r = threadIdx.x; r < oy; r += blockDim.x) {
        const double v = (double)(r + 1);
        const double x = bx + v * m.a[1][0], y = by + v * m.a[1][1], z = bz + v * m.a[1][2];
        float val = 0.f;
        if (x > -1.0 && x < (double)(nx + 2) && y > -1.0 && y < (double)(ny + 2) && z > -1.0 && z < (double)(nz + 2)) {
            const double fx0 = std::floor(x), fy0 = std::floor(y), fz0 = std::floor(z);
            const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0, iz = (int64_t)fz0;  // 1-based lower corner
            const float fx = (float)(x - fx0), fy = (float)(y - fy0), fz = (float)(z - fz0);
            float cval[2][2][2];
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx)
#pragma unroll
                    for (int dy = 0; dy < 2; ++dy) {
                        const int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
                        const bool in = jx >= 1 && jx <= nx && jy >= 1 && jy <= ny && jz >= 1 && jz <= nz;
                        cval[dz][dx][dy] = in ? vol[(jy - 1) + ny * ((jx - 1) + nx * (jz - 1))] : 0.f;
                    }
            float pl[2];
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const float a = cval[dz][0][0] + fy * (cval[dz][0][1] - cval[dz][0][0]);
                const float b = cval[dz][1][0] + fy * (cval[dz][1][1] - cval[dz][1][0]);
                pl[dz] = a + fx * (b - a);
            }
            val = pl[0] + fz * (pl[1] - pl[0]);
        }
        o[r] = val;
    }
}

void launch_affine3d(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3], const Affine3& m,
                     hipStream_t s) {
    const int64_t cols = odims[1] * odims[2];
    if (cols <= 0 || odims[0] <= 0) return;
    const unsigned nt = odims[0] >= 256 ? 256u : (unsigned)((odims[0] + 63) / 64 * 64);
    if (cols * (int64_t)nt > 0xFFFFFFFFll) fail(PK_ERR_UNSUPPORTED_SIZE, "output volume has too many columns");
    hipLaunchKernelGGL(k_affine3d, dim3((unsigned)cols), dim3(nt), 0, s, vol, out, dims[0], dims[1], dims[2], odims[0],
                       odims[1], m);
    PK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- cubic
// imwarp(..., 'cubic', 'FillValues', 0): the input padded with one voxel of
// the fill value 0 on every side, then cubic convolution (Keys, a = -0.5:
// the kernel that reproduces quadratics) over that padded grid, whose two
// missing neighbours past each end come from Keys' boundary rule
// f(-1) = 3 f(0) - 3 f(1) + f(2); a point outside the padded grid takes the
// fill value.  Per axis the (up to) four taps fold into weights on real
// samples: a tap on the zero pad drops out, the extrapolated tap below the
// grid (padded index -1, i.e. 1-based -1) adds -3 w to sample 1 and w to
// sample 2, and symmetrically at the top.
struct CubicTaps {
    int64_t idx[4];  // 0-based samples
    float w[4];
    int n;
};

__device__ __forceinline__ CubicTaps cubic_taps(double x, int64_t n) {
    // x: 1-based coordinate, within [0, n + 1]
    const double f0 = floor(x);
    const int64_t i = (int64_t)f0;
    const float t = (float)(x - f0);
    const float t2 = t * t, t3 = t2 * t;
    const float w[4] = {-0.5f * t3 + t2 - 0.5f * t, 1.5f * t3 - 2.5f * t2 + 1.f, -1.5f * t3 + 2.f * t2 + 0.5f * t,
                        0.5f * t3 - 0.5f * t2};
    CubicTaps c{};
    auto add = [&](int64_t j, float v) {  // j: 1-based real sample
        if (j < 1 || j > n) return;        // the zero pad
        for (int k = 0; k < c.n; ++k)
            if (c.idx[k] == j - 1) {
                c.w[k] += v;
                return;
            }
        c.idx[c.n] = j - 1;
        c.w[c.n] = v;
        ++c.n;
    };
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t j = i - 1 + k;
        if (w[k] == 0.f) continue;  // also the taps past the extension at x = n + 1 exactly
        if (j == -1) {               // 3 f(0) - 3 f(1) + f(2), f(0) the pad
            add(1, -3.f * w[k]);
            add(2, w[k]);
        } else if (j == n + 2) {     // 3 f(n+1) - 3 f(n) + f(n-1), f(n+1) the pad
            add(n, -3.f * w[k]);
            add(n - 1, w[k]);
        } else {
            add(j, w[k]);
        }
    }
    return c;
}

__global__ void __launch_bounds__(256) k_affine3d_cubic(const float* __restrict__ vol, float* __restrict__ out,
                                                        int64_t ny, int64_t nx, int64_t nz, int64_t oy, int64_t ox,
                                                        Affine3 m) {
    const int64_t col = blockIdx.x;
    const int64_t c = col % ox, p = col / ox;
    const double u = (double)(c + 1), w = (double)(p + 1);
    const double bx = u * m.a[0][0] + w * m.a[2][0] + m.a[3][0];
    const double by = u * m.a[0][1] + w * m.a[2][1] + m.a[3][1];
    const double bz = u * m.a[0][2] + w * m.a[2][2] + m.a[3][2];
    float* __restrict__ o = out + col * oy;
    for (int64_t r = threadIdx.x; r < oy; r += blockDim.x) {
        const double v = (double)(r + 1);
        const double x = bx + v * m.a[1][0], y = by + v * m.a[1][1], z = bz + v * m.a[1][2];
        float val = 0.f;
        if (x >= 0.0 && x <= (double)(nx + 1) && y >= 0.0 && y <= (double)(ny + 1) && z >= 0.0 &&
            z <= (double)(nz + 1)) {
            const CubicTaps tx = cubic_taps(x, nx), ty = cubic_taps(y, ny), tz = cubic_taps(z, nz);
            for (int kz = 0; kz < tz.n; ++kz) {
                float sx = 0.f;
                for (int kx = 0; kx < tx.n; ++kx) {
                    const float* __restrict__ rowp = vol + ny * (tx.idx[kx] + nx * tz.idx[kz]);
                    float sy = 0.f;
                    for (int ky = 0; ky < ty.n; ++ky) sy += ty.w[ky] * rowp[ty.idx[ky]];
                    sx += tx.w[kx] * sy;
                }
                val += tz.w[kz] * sx;
            }
        }
        o[r] = val;
    }
}

void launch_affine3d_cubic(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3],
                           const Affine3& m, hipStream_t s) {
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 2) fail(PK_ERR_INVALID_ARG, "cubic interpolation needs at least 2 voxels along every axis");
    const int64_t cols = odims[1] * odims[2];
    if (cols <= 0 || odims[0] <= 0) return;
    const unsigned nt = odims[0] >= 256 ? 256u : (unsigned)((odims[0] + 63) / 64 * 64);
    if (cols * (int64_t)nt > 0xFFFFFFFFll) fail(PK_ERR_UNSUPPORTED_SIZE, "output volume has too many columns");
    hipLaunchKernelGGL(k_affine3d_cubic, dim3((unsigned)cols), dim3(nt), 0, s, vol, out, dims[0], dims[1], dims[2],
                       odims[0], odims[1], m);
    PK_HIP(hipGetLastError());
}

}  // namespace pk

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

// pk_dsr.hip -- combined deskew and rotation of a stage-scan volume
// (microscopeDataProcessing/deskew_rotate/deskewRotateFrame3D.m:50-253, the
// reference's default after decon: RLdecon.m:22 DSRCombined = true, called at
// RLdecon.m:489-495), linear interpolation without resampling.
//
// Geometry (host, double): dx = cos(angle) dz / xyPixelSize pixels of shift
// per slice; when |dx| exceeds xStepThresh the slices are first interpolated
// in skewed space to a finer step dz' = dz / k (skewed_space_interp_defined_
// stepsize.m), so that the per-slice shift cos(angle) dz' / xyPixelSize is
// below the threshold; then one affine map -- shear by dx per slice, centre,
// z scaled by sin(angle) dz / xyPixelSize, rotation about y by the skew
// angle, re-centre on the output box -- takes output voxels back to the
// source.  For a stage scan that map keeps dim-1 rows (the source dim-2
// coordinate depends on the output dim-3 index only, the source slice on the
// output dims 2 and 3), so every output row is a bilinear blend of four whole
// source rows with fixed weights (volume_deskew_rotate_warp_mex.cpp:217-282):
// the kernels stream dim-1 rows, coalesced, HBM-bound.
//
// Two device passes: the skewed-space interpolation writes the finer-step
// volume (slices that fall on an input slice are copied), then the row warp
// reads it.  Maps that do not keep rows (objective scans) go through the
// general trilinear warp of pk_deskew.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "pk_engine.hpp"
#include "pk_misc.hpp"

namespace pk {

struct Affine3 {
    double a[4][4];  // row-vector convention: [x y z 1] = [u v w 1] * a
};
void launch_affine3d(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3], const Affine3& m,
                     hipStream_t s);
void launch_affine3d_cubic(const float* vol, float* out, const int64_t dims[3], const int64_t odims[3],
                           const Affine3& m, hipStream_t s);

struct DsrGeom {
    double tmat[4][4];  // backward map over 0-based (dim1, dim2, dim3), column-vector form
    double fwd[4][4];   // imwarp's forward matrix, 1-based [x y z 1] row vectors
    int64_t out[3];     // outSize [ny nx' nz']
    double xstep;       // |dx| of the input slices
    double stepsize;    // interpolated slice step in input slices (1: no interpolation)
    int64_t ntotal;     // slices after interpolation
    bool rows;          // the map keeps dim-1 rows
    bool resampled;     // a resampleFactor other than [1 1 1]: imwarp's form only (:73)
};

static double mround(double x) { return x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5); }

static void matmul4(const double a[4][4], const double b[4][4], double c[4][4]) {
    double t[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[i][k] * b[k][j];
            t[i][j] = s;
        }
    std::memcpy(c, t, sizeof(t));
}

// Gauss-Jordan inverse with partial pivoting (affine 4x4 matrices).
static bool invert4(const double m[4][4], double r[4][4]) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int i = c + 1; i < 4; ++i)
            if (std::fabs(a[i][c]) > std::fabs(a[p][c])) p = i;
        if (std::fabs(a[p][c]) < 1e-300) return false;
        if (p != c)
            for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
        const double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int i = 0; i < 4; ++i) {
            if (i == c) continue;
            const double f = a[i][c];
            if (f != 0)
                for (int j = 0; j < 8; ++j) a[i][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[i][j] = a[i][j + 4];
    return true;
}

DsrGeom dsr_geometry(const int64_t dims[3], double angle, double dz, double xy, bool reverse, bool objective,
                     double x_step_thresh, const double* rs = nullptr) {
    if (!(xy > 0) || !(dz > 0)) fail(PK_ERR_INVALID_ARG, "dz and xyPixelSize must be positive");
    const int64_t ny = dims[0], nx = dims[1];
    int64_t nz = dims[2];
    double theta = angle * M_PI / 180.0;
    if (!objective && std::sin(std::fabs(theta)) == 0.0) fail(PK_ERR_INVALID_ARG, "a stage scan needs a nonzero angle");
    double dx = std::cos(theta) * dz / xy;
    double zan = objective ? dz / xy : std::sin(std::fabs(theta)) * dz / xy;
    DsrGeom g{};
    // output size from the original dz (deskewRotateFrame3D.m:61-69)
    double o1, o2;
    if (!objective) {
        o1 = (double)(nx - 1) * std::cos(theta) + (double)(nz - 1) * zan / std::sin(std::fabs(theta));
        o2 = (double)(nx - 1) * std::sin(std::fabs(theta)) - 4;
    } else {
        o1 = (double)nx * std::cos(theta) + (double)nz * zan * std::sin(std::fabs(theta));
        o2 = (double)nz * zan * std::cos(theta) + (double)nx * std::sin(std::fabs(theta));
    }
    g.out[0] = ny;
    g.out[1] = (int64_t)mround(o1);
    g.out[2] = (int64_t)mround(o2);
    g.xstep = std::fabs(dx);
    g.stepsize = 1.0;
    g.ntotal = nz;
    // skewed-space interpolation step (:71-125; the fast path's slice count, :115)
    if (!objective && std::fabs(dx) > x_step_thresh) {
        double dzout;
        if (std::fabs(dx) / x_step_thresh < 1.5) {
            const double thr = xy * x_step_thresh / std::cos(theta);
            dzout = dz / (std::ceil((dz / thr) / 1.0) * 1.0);
            int counter = 1;
            while (dzout / thr < 0.4) {
                const double f = 1.0 / (double)(counter + 1);
                dzout = dz / (std::ceil((dz / thr) / f) * f);
                ++counter;
            }
        } else {
            dzout = dz / std::ceil(std::fabs(dx) / x_step_thresh);
        }
        g.stepsize = dzout / dz;
        nz = (int64_t)std::floor(std::round((double)(nz - 1) / g.stepsize * 100000.0) / 100000.0) + 1;
        g.ntotal = nz;
        dz = dzout;
        dx = std::cos(theta) * dz / xy;
        zan = std::sin(std::fabs(theta)) * dz / xy;
    }
    // deskew shear (:131-149) and rotation about y (:153-177)
    const double xshift = !reverse ? -dx : dx + std::ceil((double)(nz - 1) * dx);
    const double xstp = !reverse ? dx : -dx;
    const double nxds = objective ? (double)nx : std::ceil((double)(nz - 1) * dx) + (double)nx;
    double S_ds[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    if (!objective) {
        S_ds[2][0] = xstp;
        S_ds[3][0] = xshift;
    }
    if (reverse) theta = -theta;
    const double cx = (nxds + 1) / 2, cy = ((double)ny + 1) / 2, cz = ((double)nz + 1) / 2;
    const double ox = ((double)g.out[1] + 1) / 2, oy = ((double)g.out[0] + 1) / 2, oz = ((double)g.out[2] + 1) / 2;
    const double ct = std::cos(theta), st = std::sin(theta);
    auto chain = [&](double shift_in, double shift_out, const double ds[4][4], double M[4][4]) {
        // ds * (T1 * S * R * T2) with T1 and T2 moved by shift_in / shift_out (0-based indexing)
        double T1[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {-cx + shift_in, -cy + shift_in, -cz + shift_in, 1}};
        double S[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, zan, 0}, {0, 0, 0, 1}};
        double R[4][4] = {{ct, 0, -st, 0}, {0, 1, 0, 0}, {st, 0, ct, 0}, {0, 0, 0, 1}};
        double T2[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {ox - shift_out, oy - shift_out, oz - shift_out, 1}};
        double t[4][4];
        matmul4(T1, S, t);
        matmul4(t, R, t);
        matmul4(t, T2, t);
        matmul4(ds, t, M);
    };
    chain(0, 0, S_ds, g.fwd);  // imwarp's form (:210)
    if (rs && (rs[0] != 1.0 || rs[1] != 1.0 || rs[2] != 1.0)) {
        // resampling after deskew and rotation (:183-196): about the output
        // box's centre, RS = diag(1 ./ rs) in imwarp's x y z order, into
        // round(outSize ./ rs([2 1 3])); the forward matrix gains RT1 * RS * RT2
        for (int d = 0; d < 3; ++d)
            if (!(rs[d] > 0) || !std::isfinite(rs[d])) fail(PK_ERR_INVALID_ARG, "resampleFactor must be positive");
        const int64_t no[3] = {(int64_t)mround((double)g.out[0] / rs[1]), (int64_t)mround((double)g.out[1] / rs[0]),
                               (int64_t)mround((double)g.out[2] / rs[2])};
        if (no[0] < 1 || no[1] < 1 || no[2] < 1) fail(PK_ERR_INVALID_ARG, "resampleFactor leaves an empty output");
        const double RT1[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {-ox, -oy, -oz, 1}};
        const double RS[4][4] = {{1 / rs[0], 0, 0, 0}, {0, 1 / rs[1], 0, 0}, {0, 0, 1 / rs[2], 0}, {0, 0, 0, 1}};
        const double RT2[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0},
                                  {((double)no[1] + 1) / 2, ((double)no[0] + 1) / 2, ((double)no[2] + 1) / 2, 1}};
        double r[4][4];
        matmul4(RT1, RS, r);
        matmul4(r, RT2, r);
        matmul4(g.fwd, r, g.fwd);
        for (int d = 0; d < 3; ++d) g.out[d] = no[d];
        g.resampled = true;
    }
    double ds2[4][4];
    std::memcpy(ds2, S_ds, sizeof(ds2));
    if (!objective && reverse) ds2[3][0] -= dx;  // :215-217
    double M[4][4], Mt[4][4], inv[4][4];
    chain(1, 1, ds2, M);  // (T1 + offset), (T2 - offset): 0-based coordinates (:212-218)
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Mt[i][j] = M[j][i];
    if (!invert4(Mt, inv)) fail(PK_ERR_INVALID_ARG, "singular deskew / rotation transform");
    static const int perm[4] = {1, 0, 2, 3};  // (x, y, z) -> (dim1, dim2, dim3)
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g.tmat[i][j] = inv[perm[i]][perm[j]];
    const double tol = 1e-9;
    const double (&t)[4][4] = g.tmat;
    g.rows = std::fabs(t[0][0] - 1) < tol && std::fabs(t[0][1]) < tol && std::fabs(t[0][2]) < tol &&
             std::fabs(t[0][3]) < tol && std::fabs(t[1][0]) < tol && std::fabs(t[1][1]) < tol &&
             std::fabs(t[2][0]) < tol;
    return g;
}

#pragma clang fp contract(off)

template <class T>
__device__ __forceinline__ float ld(const T* p, int64_t i) {
    return (float)p[i];
}

// skewed_space_interp_defined_stepsize.m:29-100: slice zi of the finer-step
// volume, zout = 1 + zi * stepsize (1-based), blends slice floor(zout) sampled
// x + frac(zout) * xstep (reverse: x - frac * xstep) along dim 2 and the next
// slice at x - (1 - frac) * xstep (reverse: x + (1 - frac) * xstep), each by
// linear interpolation over zero padding; a slice on an input slice is copied.
// One thread per dim-1 element: rows stream coalesced.
template <class T>
__global__ void __launch_bounds__(256) k_skewed_interp(const T* __restrict__ in, float* __restrict__ out, int64_t ny,
                                                       int64_t nx, int64_t nz, double xstep, double stepsize,
                                                       int reverse) {
    const int64_t row = blockIdx.y;  // dim-2 index j and output slice zi
    const int64_t j = row % nx, zi = row / nx;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ny) return;
    const double zout = 1.0 + (double)zi * stepsize;
    const int64_t zs = (int64_t)std::floor(zout);
    const double zw = zout - (double)zs;
    const int64_t plane = ny * nx;
    float* __restrict__ o = out + zi * plane + j * ny + i;
    if (zs >= nz) {
        *o = ld(in, (nz - 1) * plane + j * ny + i);
        return;
    }
    const double xa = std::ceil(xstep);
    double szm, tzm;
    if (reverse) {
        szm = xa - zw * xstep + 1;
        tzm = szm - (xa - xstep);
    } else {
        szm = zw * xstep + 1;
        tzm = szm + (xa - xstep);
    }
    const double s = std::floor(szm), t = std::floor(tzm);
    const double sw = szm - s, tw = tzm - t;
    const T* __restrict__ a = in + (zs - 1) * plane;  // slice floor(zout)
    if ((reverse && sw == 0) || (!reverse && tw == 0)) {
        *o = ld(a, j * ny + i);
        return;
    }
    const T* __restrict__ b = in + zs * plane;  // the next slice
    // padded column k (0-based) of the s / t arrays -> source column, or -1 for the zero padding
    const int64_t ia = (int64_t)xa;
    auto col_s = [&](int64_t k) { return reverse ? k - ia : k; };
    auto col_t = [&](int64_t k) { return reverse ? k : k - ia; };
    auto val = [&](const T* __restrict__ v, int64_t c) { return (c >= 0 && c < nx) ? ld(v, c * ny + i) : 0.f; };
    const int64_t ks = (int64_t)s - 1 + j, kt = (int64_t)t - 1 + j;
    const float sq = val(a, col_s(ks)) * (float)(1 - sw) + val(a, col_s(ks + 1)) * (float)sw;
    const float tq = val(b, col_t(kt)) * (float)(1 - tw) + val(b, col_t(kt + 1)) * (float)tw;
    *o = sq * (float)(1 - zw) + tq * (float)zw;
}

// volume_deskew_rotate_warp_mex.cpp:217-282: output row (y, z) of the box
// starting at (bx, by, bz) reads source rows around (ty, tz), ty = t12 (z + bz)
// + t13, tz = t21 (y + by) + t22 (z + bz) + t23 (0-based), blended
// ((a w00 + b w01) + c w10) + d w11 from rows (ty0, tz0), (ty0, tz0 + 1),
// (ty0 + 1, tz0), (ty0 + 1, tz0 + 1); rows with a coordinate outside
// [0, n - 1) are zero, a coordinate of exactly n - 1 reads the last row.
struct RowMap {
    double t12, t13, t21, t22, t23;
};

template <class T, class O>
__global__ void __launch_bounds__(256) k_dsr_rows(const T* __restrict__ in, O* __restrict__ out, int64_t sx,
                                                  int64_t sy, int64_t sz, int64_t ox, int64_t oy, int64_t bx,
                                                  int64_t by, int64_t bz, RowMap m) {
    const int64_t row = blockIdx.y;
    const int64_t y = row % oy, z = row / oy;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ox) return;
    O* __restrict__ o = out + row * ox + x;
    auto put = [&](float v) {
        if constexpr (sizeof(O) == 2) *o = (O)u16_round(v);
        else *o = v;
    };
    double ty = m.t12 * (double)(z + bz) + m.t13;
    double tz = m.t21 * (double)(y + by) + m.t22 * (double)(z + bz) + m.t23;
    int64_t y0, z0;
    float wy, wz;
    if (ty >= 0 && ty + 1 < (double)sy) {
        y0 = (int64_t)std::floor(ty);
        wy = (float)(ty - (double)y0);
    } else if (ty + 1 == (double)sy && sy >= 2) {
        y0 = sy - 2;
        wy = 1.f;
    } else {
        put(0.f);
        return;
    }
    if (tz >= 0 && tz + 1 < (double)sz) {
        z0 = (int64_t)std::floor(tz);
        wz = (float)(tz - (double)z0);
    } else if (tz + 1 == (double)sz && sz >= 2) {
        z0 = sz - 2;
        wz = 1.f;
    } else {
        put(0.f);
        return;
    }
    const float w00 = (1 - wy) * (1 - wz), w01 = (1 - wy) * wz, w10 = wy * (1 - wz), w11 = wy * wz;
    const int64_t plane = sx * sy;
    const int64_t base = (bx + x) + y0 * sx + z0 * plane;
    const float a = ld(in, base), b = ld(in, base + plane), c = ld(in, base + sx), d = ld(in, base + sx + plane);
    put(((a * w00 + b * w01) + c * w10) + d * w11);
}

__global__ void k_u16_to_f32(const uint16_t* __restrict__ in, float* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (float)in[i];
}

// in place: the value uint16() of the single value would hold (the skewed-space
// interpolation's uint16 output, which the cubic imwarp then reads)
__global__ void k_round_u16(float* __restrict__ v, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (float)u16_round(v[i]);
}

__global__ void k_f32_to_u16(const float* __restrict__ in, uint16_t* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint16_t)u16_round(in[i]);
}

#pragma clang fp contract(on)

void dsr_output_size(const int64_t dims[3], double angle, double dz, double xy, bool reverse, bool objective,
                     double x_step_thresh, int64_t out[3], const double* rs) {
    const DsrGeom g = dsr_geometry(dims, angle, dz, xy, reverse, objective, x_step_thresh, rs);
    for (int d = 0; d < 3; ++d) out[d] = g.out[d];
}

static unsigned blocks_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

void pk_ctx_impl::deskew_rotate(const pk_vol* in, double angle, double dz, double xy, bool reverse, bool objective,
                                double x_step_thresh, const int64_t* bbox, pk_vol* out, int interp, const double* rs,
                                bool imwarp_form) {
    if (!in || !out || !in->data || !out->data) fail(PK_ERR_INVALID_ARG, "null volume");
    if (in->mem != PK_MEM_DEVICE || out->mem != PK_MEM_DEVICE) fail(PK_ERR_INVALID_ARG, "device volumes expected");
    if (in->dtype != PK_UINT16 && in->dtype != PK_SINGLE) fail(PK_ERR_UNSUPPORTED_TYPE, "input class uint16 or single");
    if (out->dtype != PK_UINT16 && out->dtype != PK_SINGLE) fail(PK_ERR_UNSUPPORTED_TYPE, "output class uint16 or single");
    for (int d = 0; d < 3; ++d)
        if (in->dims[d] < 2) fail(PK_ERR_INVALID_ARG, "every axis needs at least 2 voxels");
    const DsrGeom g = dsr_geometry(in->dims, angle, dz, xy, reverse, objective, x_step_thresh, rs);
    int64_t b[6] = {1, 1, 1, g.out[0], g.out[1], g.out[2]};
    if (bbox) {
        std::memcpy(b, bbox, sizeof(b));
        for (int d = 0; d < 3; ++d)
            if (b[d] < 1 || b[d + 3] < b[d]) fail(PK_ERR_INVALID_ARG, "bbox: 1-based start <= end expected");
    }
    // the fast path unless the reference leaves it (:72-73: gpuProcess, a
    // resampling factor, a method other than linear) or the map does not keep rows
    const bool leave_fast = imwarp_form || interp == 1 || g.resampled;
    const bool general = !g.rows || leave_fast;
    // the row-by-row warp maps output rows to input rows one to one
    if (!general && b[3] > in->dims[0]) fail(PK_ERR_INVALID_ARG, "bbox: dim-1 range beyond the volume's rows");
    const int64_t od[3] = {b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1};
    for (int d = 0; d < 3; ++d)
        if (out->dims[d] != od[d]) fail(PK_ERR_INVALID_ARG, "output dims must equal the output box");
    if (od[0] < 1 || od[1] < 1 || od[2] < 1) fail(PK_ERR_INVALID_ARG, "empty output");
    const int64_t ny = in->dims[0], nx = in->dims[1], nz = in->dims[2];
    const int64_t on = od[0] * od[1] * od[2];
    const bool u16in = in->dtype == PK_UINT16, u16out = out->dtype == PK_UINT16;
    if (general) {
        // the general imwarp form (:210; off the fast path, :73 and :98-110,
        // the skewed-space interpolation first when the step is large), linear
        // or cubic over zero padding; a bbox is imref3d's output window
        // (:203-205), an offset of the output grid
        const bool skew = leave_fast && g.stepsize != 1.0;
        const int64_t sz = skew ? g.ntotal : nz;
        const float* src = static_cast<const float*>(in->data);
        size_t need = 0;
        if (u16in || skew) need += (sizeof(float) * (size_t)(ny * nx * sz) + 255) / 256 * 256;
        if (u16out) need += sizeof(float) * (size_t)on + 256;
        ensure(ws.dsr, need);
        char* p = ws.dsr.as<char>();
        if (skew) {
            float* f = reinterpret_cast<float*>(p);
            const dim3 grid((unsigned)((ny + 255) / 256), (unsigned)(nx * sz));
            if (nx * sz > 0x7FFFFFFF) fail(PK_ERR_UNSUPPORTED_SIZE, "too many rows");
            if (u16in)
                hipLaunchKernelGGL(k_skewed_interp<uint16_t>, grid, dim3(256), 0, stream,
                                   static_cast<const uint16_t*>(in->data), f, ny, nx, nz, g.xstep, g.stepsize,
                                   reverse ? 1 : 0);
            else
                hipLaunchKernelGGL(k_skewed_interp<float>, grid, dim3(256), 0, stream,
                                   static_cast<const float*>(in->data), f, ny, nx, nz, g.xstep, g.stepsize,
                                   reverse ? 1 : 0);
            // the reference's interpolation returns uint16 for a uint16 volume with save16bit
            // (:105), and the cubic imwarp reads that
            if (u16in && u16out)
                hipLaunchKernelGGL(k_round_u16, dim3(blocks_for(ny * nx * sz)), dim3(256), 0, stream, f,
                                   ny * nx * sz);
            src = f;
            p += (sizeof(float) * (size_t)(ny * nx * sz) + 255) / 256 * 256;
        } else if (u16in) {
            float* f = reinterpret_cast<float*>(p);
            hipLaunchKernelGGL(k_u16_to_f32, dim3(blocks_for(ny * nx * nz)), dim3(256), 0, stream,
                               static_cast<const uint16_t*>(in->data), f, ny * nx * nz);
            src = f;
            p += (sizeof(float) * (size_t)(ny * nx * nz) + 255) / 256 * 256;
        }
        float* dst = u16out ? reinterpret_cast<float*>(p) : static_cast<float*>(out->data);
        Affine3 m, mi;
        if (!invert4(g.fwd, mi.a)) fail(PK_ERR_INVALID_ARG, "singular transform");
        // output voxel (r, c, p) of the window is world point (c + b2 - 1, r + b1 - 1, p + b3 - 1)
        const double off[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0},
                                  {(double)(b[1] - 1), (double)(b[0] - 1), (double)(b[2] - 1), 1}};
        matmul4(off, mi.a, m.a);
        const int64_t sd[3] = {ny, nx, sz};
        if (interp == 1)
            launch_affine3d_cubic(src, dst, sd, od, m, stream);
        else
            launch_affine3d(src, dst, sd, od, m, stream);
        if (u16out)
            hipLaunchKernelGGL(k_f32_to_u16, dim3(blocks_for(on)), dim3(256), 0, stream, dst,
                               static_cast<uint16_t*>(out->data), on);
        PK_HIP(hipGetLastError());
        PK_HIP(hipStreamSynchronize(stream));
        return;
    }
    const void* src = in->data;
    bool srcf = !u16in;
    int64_t sz = nz;
    if (g.stepsize != 1.0) {  // skewed-space interpolation into the workspace
        sz = g.ntotal;
        ensure(ws.dsr, sizeof(float) * (size_t)(ny * nx * sz));
        float* f = ws.dsr.as<float>();
        const dim3 grid((unsigned)((ny + 255) / 256), (unsigned)(nx * sz));
        if (nx * sz > 0x7FFFFFFF) fail(PK_ERR_UNSUPPORTED_SIZE, "too many rows");
        if (u16in)
            hipLaunchKernelGGL(k_skewed_interp<uint16_t>, grid, dim3(256), 0, stream,
                               static_cast<const uint16_t*>(in->data), f, ny, nx, nz, g.xstep, g.stepsize, reverse ? 1 : 0);
        else
            hipLaunchKernelGGL(k_skewed_interp<float>, grid, dim3(256), 0, stream, static_cast<const float*>(in->data),
                               f, ny, nx, nz, g.xstep, g.stepsize, reverse ? 1 : 0);
        PK_HIP(hipGetLastError());
        src = f;
        srcf = true;
    }
    const RowMap m{g.tmat[1][2], g.tmat[1][3], g.tmat[2][1], g.tmat[2][2], g.tmat[2][3]};
    if (od[1] * od[2] > 0x7FFFFFFF) fail(PK_ERR_UNSUPPORTED_SIZE, "too many output rows");
    const dim3 grid((unsigned)((od[0] + 255) / 256), (unsigned)(od[1] * od[2]));
    const int64_t bx = b[0] - 1, by = b[1] - 1, bz = b[2] - 1;
    if (srcf) {
        const float* s = static_cast<const float*>(src);
        if (u16out)
            hipLaunchKernelGGL((k_dsr_rows<float, uint16_t>), grid, dim3(256), 0, stream, s,
                               static_cast<uint16_t*>(out->data), ny, nx, sz, od[0], od[1], bx, by, bz, m);
        else
            hipLaunchKernelGGL((k_dsr_rows<float, float>), grid, dim3(256), 0, stream, s, static_cast<float*>(out->data),
                               ny, nx, sz, od[0], od[1], bx, by, bz, m);
    } else {
        const uint16_t* s = static_cast<const uint16_t*>(src);
        if (u16out)
            hipLaunchKernelGGL((k_dsr_rows<uint16_t, uint16_t>), grid, dim3(256), 0, stream, s,
                               static_cast<uint16_t*>(out->data), ny, nx, sz, od[0], od[1], bx, by, bz, m);
        else
            hipLaunchKernelGGL((k_dsr_rows<uint16_t, float>), grid, dim3(256), 0, stream, s,
                               static_cast<float*>(out->data), ny, nx, sz, od[0], od[1], bx, by, bz, m);
    }
    PK_HIP(hipGetLastError());
    PK_HIP(hipStreamSynchronize(stream));
}

}  // namespace pk

// pk_probe.hip -- HBM ceilings of the deconvolution passes' access patterns
// (diagnostics; no counterpart in the reference).
//
// Each probe copies one pitched half spectrum (the layout every pass works on:
// hp >= h1 = n1/2 + 1 complex points per row, n2 rows, n3 planes) into a
// second buffer with the access pattern of one pass and no arithmetic:
//   - sequential: 16-byte loads and stores over the whole buffer;
//   - dim-2 columns (the y passes): tiles of T adjacent columns, column
//     stride hp, one tile row per z plane;
//   - dim-3 columns (z_conv's transforms): tiles of T adjacent columns,
//     column stride hp * n2 (the pad columns hp - h1 skipped, as the passes
//     skip them);
//   - dim-3 columns with a second input read at the same points (z_conv's
//     traffic: spectrum and OTF in, spectrum out; the values are added so
//     the second read is kept).
// A column tile is moved as the passes move it: block = T columns x P
// threads, thread (c, p) holds the E points at rows p + e P of its column in
// registers, all loads before all stores, with the strided passes' cache
// policy (PK_NT_COL: nontemporal stores for tiles of 16 or more columns, the
// second input loaded as the convolution loads its OTF).  The block-to-tile order
// can be the passes' XCD-aware one.  The sequential copy is a plain copy
// (temporal loads and stores).  Timed with HIP events on the context's stream:
// back to back (a copy may then read what the previous one left in the
// memory-side cache), or cold, each copy after a read of another 1 GiB
// buffer, as a pass meets its input after the passes before it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pk_engine.hpp"

namespace pk {

namespace {

template <int E>
__global__ void __launch_bounds__(1024) k_probe_cols(const float2* __restrict__ src, const float2* __restrict__ src2,
                                                     float2* __restrict__ dst, int64_t acount, int64_t amod,
                                                     int64_t alim, int64_t bstride, int64_t stride, int T, int P,
                                                     int xcd) {
    int64_t bx = blockIdx.x, by = blockIdx.y;
    if (xcd) {  // XCD k walks the k-th eighth of the tiles, in order (the passes' remap)
        const int64_t gx = gridDim.x, total = gx * gridDim.y;
        const int64_t b = bx + by * gx;
        const int64_t t = (b & 7) * (total >> 3) + (b >> 3);
        bx = t % gx;
        by = t / gx;
    }
    const int t = threadIdx.x;
    const int p = t / T;
    const int c = t - p * T;
    const int64_t col = bx * T + c;
    if (p >= P || col >= acount || (amod && col % amod >= alim)) return;  // no barrier below
    const int64_t step = (int64_t)P * stride;
    const int64_t off0 = col + by * bstride + (int64_t)p * stride;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = src[off0 + e * step];
    if (src2) {
        float2 h[E];
#pragma unroll
        for (int e = 0; e < E; ++e) h[e] = ld_f2<kNtOtfLd>(src2 + off0 + e * step);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = make_float2(v[e].x + h[e].x, v[e].y + h[e].y);
    }
    if (kNtColSt && T >= 16) {
#pragma unroll
        for (int e = 0; e < E; ++e) st_f2<true>(dst + off0 + e * step, v[e]);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) st_f2<false>(dst + off0 + e * step, v[e]);
    }
}

// Reads n float4 (the cold probes' cache flush: read-only, so no dirty lines
// are left to write back during the timed copy); the sum is stored only under
// a condition no finite input meets, which keeps the loads.
__global__ void k_probe_flush(const float4* __restrict__ src, int64_t n, float* __restrict__ sink) {
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = src[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == -1.2345e-38f) sink[0] = s;
}

__global__ void k_probe_seq(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <int E>
void launch_cols(dim3 grid, dim3 block, hipStream_t st, const float2* s, const float2* s2, float2* d, int64_t acount,
                 int64_t amod, int64_t alim, int64_t bstride, int64_t stride, int T, int P, int xcd) {
    hipLaunchKernelGGL(k_probe_cols<E>, grid, block, 0, st, s, s2, d, acount, amod, alim, bstride, stride, T, P, xcd);
}

}  // namespace

double pk_ctx_impl::probe_copy(int pattern, const int64_t dims[3], int tile, int vals, int flags, int reps,
                               double* bytes) {
    int xcd = flags & 1;
    const bool cold = (flags & 2) != 0;
    const VolPlan& vp = plan(dims[0], dims[1], dims[2]);
    ensure(ws.A, sizeof(float2) * (size_t)vp.S);
    ensure(ws.tmp, sizeof(float2) * (size_t)vp.S);
    const float2* src = ws.A.as<float2>();
    float2* dst = ws.tmp.as<float2>();
    PK_HIP(hipMemsetAsync(ws.A.p, 0, sizeof(float2) * (size_t)vp.S, stream));
    const float2* src2 = nullptr;
    if (pattern == 3) {  // the second input: a third buffer of the same layout
        ensure(ws.J3, sizeof(float2) * (size_t)vp.S);
        PK_HIP(hipMemsetAsync(ws.J3.p, 0, sizeof(float2) * (size_t)vp.S, stream));
        src2 = ws.J3.as<float2>();
    }
    dim3 grid, block;
    int64_t acount = 0, amod = 0, alim = 0, bstride = 0, stride = 0;
    int P = 0;
    if (pattern == 0) {
        if (vp.S % 2) fail(PK_ERR_INVALID_ARG, "sequential probe needs an even number of points");
        *bytes = 2.0 * 8.0 * (double)vp.S;
        grid = dim3(2048);
        block = dim3(256);
    } else {
        const int64_t L = pattern == 1 ? vp.n2 : vp.n3;  // patterns 2 and 3: dim-3 columns
        if (vals != 4 && vals != 8 && vals != 16 && vals != 32) fail(PK_ERR_INVALID_ARG, "values per thread: 4, 8, 16, 32");
        if (L % vals) fail(PK_ERR_INVALID_ARG, "column length must be a multiple of the values per thread");
        P = (int)(L / vals);
        if (tile < 1 || (int64_t)tile * P > 1024) fail(PK_ERR_INVALID_ARG, "tile x threads per column above 1024");
        block = dim3((unsigned)(tile * P));
        if (pattern == 1) {
            acount = vp.h1;
            bstride = vp.hp * vp.n2;
            stride = vp.hp;
            grid = dim3((unsigned)((vp.h1 + tile - 1) / tile), (unsigned)vp.n3);
        } else {
            acount = vp.hp * vp.n2;
            amod = vp.hp;
            alim = vp.h1;
            stride = vp.hp * vp.n2;
            grid = dim3((unsigned)((acount + tile - 1) / tile));
        }
        if (xcd && ((int64_t)grid.x * grid.y) % 8) xcd = 0;
        *bytes = (pattern == 3 ? 3.0 : 2.0) * 8.0 * (double)vp.Sa;
    }
    auto launch = [&] {
        if (pattern == 0) {
            hipLaunchKernelGGL(k_probe_seq, grid, block, 0, stream, reinterpret_cast<const float4*>(src),
                               reinterpret_cast<float4*>(dst), vp.S / 2);
        } else if (vals == 4) {
            launch_cols<4>(grid, block, stream, src, src2, dst, acount, amod, alim, bstride, stride, tile, P, xcd);
        } else if (vals == 8) {
            launch_cols<8>(grid, block, stream, src, src2, dst, acount, amod, alim, bstride, stride, tile, P, xcd);
        } else if (vals == 16) {
            launch_cols<16>(grid, block, stream, src, src2, dst, acount, amod, alim, bstride, stride, tile, P, xcd);
        } else {
            launch_cols<32>(grid, block, stream, src, src2, dst, acount, amod, alim, bstride, stride, tile, P, xcd);
        }
        PK_HIP(hipGetLastError());
    };
    launch();  // warm-up
    // cold copies: a buffer four times the memory-side cache is read before
    // each copy (outside the timed interval)
    constexpr size_t kFlushBytes = size_t(1) << 30;
    struct DevBuf {
        void* p = nullptr;
        ~DevBuf() {
            if (p) (void)hipFree(p);
        }
    } flush;
    if (cold) {
        PK_HIP(hipMalloc(&flush.p, kFlushBytes + 256));
        PK_HIP(hipMemsetAsync(flush.p, 0, kFlushBytes + 256, stream));
    }
    hipEvent_t e0, e1;
    PK_HIP(hipEventCreate(&e0));
    PK_HIP(hipEventCreate(&e1));
    double total = 0.0;
    const int rounds = cold ? reps : 1;
    for (int k = 0; k < rounds; ++k) {
        if (cold) {
            hipLaunchKernelGGL(k_probe_flush, dim3(2048), dim3(256), 0, stream,
                               static_cast<const float4*>(flush.p), (int64_t)(kFlushBytes / sizeof(float4)),
                               reinterpret_cast<float*>(static_cast<char*>(flush.p) + kFlushBytes));
            PK_HIP(hipGetLastError());
        }
        PK_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < (cold ? 1 : reps); ++r) launch();
        PK_HIP(hipEventRecord(e1, stream));
        PK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PK_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return total / reps;
}

}  // namespace pk

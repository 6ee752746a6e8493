// pk_launch.hpp -- host-side launchers of the templated transform kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pk_fft_dev.hpp"

// Values-per-thread counts compiled into the transform kernels (the split
// across pk_kern_{x,col}_*.hip must cover the same list).
#define PK_E_LIST(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32)

namespace pk {
template <int E>
inline void launch_modes_x(dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a, const FftPlan& pl) {
    switch (a.mode) {
        case XM_R2C: hipLaunchKernelGGL((k_xpass<E, XM_R2C>), grid, block, lds, s, a, pl); break;
        case XM_C2R: hipLaunchKernelGGL((k_xpass<E, XM_C2R>), grid, block, lds, s, a, pl); break;
        case XM_PREP_U16: hipLaunchKernelGGL((k_xpass<E, XM_PREP_U16>), grid, block, lds, s, a, pl); break;
        case XM_PREP_F32: hipLaunchKernelGGL((k_xpass<E, XM_PREP_F32>), grid, block, lds, s, a, pl); break;
        case XM_RATIO: hipLaunchKernelGGL((k_xpass<E, XM_RATIO>), grid, block, lds, s, a, pl); break;
        case XM_OMW_UPDATE: hipLaunchKernelGGL((k_xpass<E, XM_OMW_UPDATE>), grid, block, lds, s, a, pl); break;
        case XM_BIGGS_Y: hipLaunchKernelGGL((k_xpass<E, XM_BIGGS_Y>), grid, block, lds, s, a, pl); break;
        case XM_BIGGS_UPDATE: hipLaunchKernelGGL((k_xpass<E, XM_BIGGS_UPDATE>), grid, block, lds, s, a, pl); break;
        case XM_OMW_FINAL: hipLaunchKernelGGL((k_xpass<E, XM_OMW_FINAL>), grid, block, lds, s, a, pl); break;
        case XM_BIGGS_FINAL: hipLaunchKernelGGL((k_xpass<E, XM_BIGGS_FINAL>), grid, block, lds, s, a, pl); break;
        default: break;
    }
}
// Dynamic LDS above 64 KB needs the kernel's limit raised once (gfx950: 160 KB
// per workgroup).
constexpr size_t kMaxLds = 160 * 1024;
// Raised once per kernel (its first launch happens outside any graph capture).
void raise_lds_once(const void* kernel);
template <class K>
inline void allow_large_lds(K kernel) {
    raise_lds_once(reinterpret_cast<const void*>(kernel));
}
template <int E, int NT = kMaxThreads, bool MIXED = false>
inline void launch_modes_col(dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a, const FftPlan& pl) {
    auto go = [&](auto kernel) {
        if (lds > 65536) allow_large_lds(kernel);
        hipLaunchKernelGGL(kernel, grid, block, lds, s, a, pl);
    };
    switch (a.mode) {
        case CM_FWD: go(k_colpass<E, CM_FWD, NT, MIXED>); break;
        case CM_INV: go(k_colpass<E, CM_INV, NT, MIXED>); break;
        case CM_CONV: go(k_colpass<E, CM_CONV, NT, MIXED>); break;
        case CM_CONVJ: go(k_colpass<E, CM_CONVJ, NT, MIXED>); break;
        default: break;
    }
}

hipError_t launch_xpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
hipError_t launch_xpass_batch(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                              const int64_t* starts, int n);
hipError_t launch_colpass_batch(int E, int nt, bool mixed, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode,
                                const ColChunk* ch, const int64_t* starts, int n);
hipError_t launch_colpass(const FftPlan& pl, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                          const ColArgs& a);
bool e_supported(int E);
template <int L, int E, bool ODD = false>
inline void launch_modes_xrow(dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    const dim3 b(kMaxThreads);
    switch (a.mode) {
        case XM_R2C: hipLaunchKernelGGL((k_xrow<L, E, XM_R2C, ODD>), grid, b, lds, s, a, tw); break;
        case XM_C2R: hipLaunchKernelGGL((k_xrow<L, E, XM_C2R, ODD>), grid, b, lds, s, a, tw); break;
        case XM_PREP_U16: hipLaunchKernelGGL((k_xrow<L, E, XM_PREP_U16, ODD>), grid, b, lds, s, a, tw); break;
        case XM_PREP_F32: hipLaunchKernelGGL((k_xrow<L, E, XM_PREP_F32, ODD>), grid, b, lds, s, a, tw); break;
        case XM_RATIO: hipLaunchKernelGGL((k_xrow<L, E, XM_RATIO, ODD>), grid, b, lds, s, a, tw); break;
        case XM_OMW_UPDATE: hipLaunchKernelGGL((k_xrow<L, E, XM_OMW_UPDATE, ODD>), grid, b, lds, s, a, tw); break;
        case XM_BIGGS_Y: hipLaunchKernelGGL((k_xrow<L, E, XM_BIGGS_Y, ODD>), grid, b, lds, s, a, tw); break;
        case XM_BIGGS_UPDATE: hipLaunchKernelGGL((k_xrow<L, E, XM_BIGGS_UPDATE, ODD>), grid, b, lds, s, a, tw); break;
        case XM_OMW_FINAL: hipLaunchKernelGGL((k_xrow<L, E, XM_OMW_FINAL, ODD>), grid, b, lds, s, a, tw); break;
        case XM_RATIO_U16: hipLaunchKernelGGL((k_xrow<L, E, XM_RATIO_U16, ODD>), grid, b, lds, s, a, tw); break;
        case XM_OMW_UPDATE_U16:
            hipLaunchKernelGGL((k_xrow<L, E, XM_OMW_UPDATE_U16, ODD>), grid, b, lds, s, a, tw);
            break;
        case XM_BIGGS_FINAL: hipLaunchKernelGGL((k_xrow<L, E, XM_BIGGS_FINAL, ODD>), grid, b, lds, s, a, tw); break;
        default: break;
    }
}
template <int L, int E, bool ODD = false>
inline void launch_modes_xrowb(dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts,
                               int n) {
    const dim3 b(kMaxThreads);
    switch (mode) {
        case XM_R2C: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_R2C, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_C2R: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_C2R, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_PREP_U16: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_PREP_U16, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_PREP_F32: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_PREP_F32, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_RATIO: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_RATIO, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_OMW_UPDATE: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_OMW_UPDATE, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_RATIO_U16:
            hipLaunchKernelGGL((k_xrow_batch<L, E, XM_RATIO_U16, ODD>), grid, b, lds, s, ch, starts, n);
            break;
        case XM_OMW_UPDATE_U16:
            hipLaunchKernelGGL((k_xrow_batch<L, E, XM_OMW_UPDATE_U16, ODD>), grid, b, lds, s, ch, starts, n);
            break;
        case XM_BIGGS_Y: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_BIGGS_Y, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_BIGGS_UPDATE:
            hipLaunchKernelGGL((k_xrow_batch<L, E, XM_BIGGS_UPDATE, ODD>), grid, b, lds, s, ch, starts, n);
            break;
        case XM_OMW_FINAL: hipLaunchKernelGGL((k_xrow_batch<L, E, XM_OMW_FINAL, ODD>), grid, b, lds, s, ch, starts, n); break;
        case XM_BIGGS_FINAL:
            hipLaunchKernelGGL((k_xrow_batch<L, E, XM_BIGGS_FINAL, ODD>), grid, b, lds, s, ch, starts, n);
            break;
        default: break;
    }
}
// Batched compile-time row kernels (same (L, E) pairs); false when absent.
bool launch_xrow_batch(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                       const int64_t* starts, int n);
// Compile-time row kernels for the (L, E) pairs in pk_kern_xrow_*.hip; false
// when the pair has none (the caller then uses the generic kernel).  L is the
// transform length: n1 / 2 for even rows, -n1 for odd rows (ODD kernels).
bool launch_xrow(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool xrow_available(int L, int E);
template <int E>
inline void launch_modes_xb(dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                            const int64_t* starts, int n) {
    switch (mode) {
        case XM_R2C: hipLaunchKernelGGL((k_xpass_batch<E, XM_R2C>), grid, block, lds, s, ch, starts, n); break;
        case XM_PREP_U16: hipLaunchKernelGGL((k_xpass_batch<E, XM_PREP_U16>), grid, block, lds, s, ch, starts, n); break;
        case XM_PREP_F32: hipLaunchKernelGGL((k_xpass_batch<E, XM_PREP_F32>), grid, block, lds, s, ch, starts, n); break;
        case XM_RATIO: hipLaunchKernelGGL((k_xpass_batch<E, XM_RATIO>), grid, block, lds, s, ch, starts, n); break;
        case XM_OMW_UPDATE: hipLaunchKernelGGL((k_xpass_batch<E, XM_OMW_UPDATE>), grid, block, lds, s, ch, starts, n); break;
        case XM_BIGGS_Y: hipLaunchKernelGGL((k_xpass_batch<E, XM_BIGGS_Y>), grid, block, lds, s, ch, starts, n); break;
        case XM_BIGGS_UPDATE: hipLaunchKernelGGL((k_xpass_batch<E, XM_BIGGS_UPDATE>), grid, block, lds, s, ch, starts, n); break;
        case XM_OMW_FINAL: hipLaunchKernelGGL((k_xpass_batch<E, XM_OMW_FINAL>), grid, block, lds, s, ch, starts, n); break;
        case XM_BIGGS_FINAL: hipLaunchKernelGGL((k_xpass_batch<E, XM_BIGGS_FINAL>), grid, block, lds, s, ch, starts, n); break;
        default: break;
    }
}
template <int E, int NT = kMaxThreads, bool MIXED = false>
inline void launch_modes_colb(dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                              const int64_t* starts, int n) {
    auto go = [&](auto kernel) {
        if (lds > 65536) allow_large_lds(kernel);
        hipLaunchKernelGGL(kernel, grid, block, lds, s, ch, starts, n);
    };
    switch (mode) {
        case CM_FWD: go(k_colpass_batch<E, CM_FWD, NT, MIXED>); break;
        case CM_INV: go(k_colpass_batch<E, CM_INV, NT, MIXED>); break;
        case CM_CONV: go(k_colpass_batch<E, CM_CONV, NT, MIXED>); break;
        case CM_CONVJ: go(k_colpass_batch<E, CM_CONVJ, NT, MIXED>); break;
        default: break;
    }
}
// Column kernels for plans with LDS-fed stages (pk_kern_col_mixed.hip):
// 256-thread blocks, values-per-thread counts listed by col_mixed_available.
bool col_mixed_available(int E, int nt);
bool launch_col_mixed(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_mixed(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                       const int64_t* starts, int n);
// Column kernels compiled for 512- / 1024-thread blocks (pk_kern_col_wide.hip).
bool col_wide_available(int E, int nt);
bool launch_col_wide(const FftPlan& pl, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_wide(int E, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                      const int64_t* starts, int n);
}  // namespace pk

// pk_launch.hip -- dispatch of a transform plan to the kernel instantiation
// compiled for its values-per-thread count.
#include "pk_launch.hpp"

#include <mutex>
#include <set>

namespace pk {
void raise_lds_once(const void* kernel) {
    static std::mutex mu;
    static std::set<const void*> seen;
    std::lock_guard<std::mutex> lk(mu);
    if (seen.insert(kernel).second)
        (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
}

bool launch_x_part_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_x_part_b(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_b(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_x_part_c(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_c(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_x_part_d(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_d(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_x_part_e(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_e(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_x_part_f(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
bool launch_xb_part_f(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch, const int64_t* starts, int n);
bool launch_col_part_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_part_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch, const int64_t* starts, int n);
bool launch_col_part_b(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_part_b(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch, const int64_t* starts, int n);
bool launch_col_part_c(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_part_c(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch, const int64_t* starts, int n);

hipError_t launch_xpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    const bool ok = launch_x_part_a(pl, grid, block, lds, s, a) ||
           launch_x_part_b(pl, grid, block, lds, s, a) ||
           launch_x_part_c(pl, grid, block, lds, s, a) ||
           launch_x_part_d(pl, grid, block, lds, s, a) ||
           launch_x_part_e(pl, grid, block, lds, s, a) ||
           launch_x_part_f(pl, grid, block, lds, s, a);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_colpass(const FftPlan& pl, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                          const ColArgs& a) {
    if (plan_mixed_stages(pl)) {
        if (nt != kMaxThreads || !launch_col_mixed(pl, grid, block, lds, s, a)) return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (nt != kMaxThreads) {
        if (!launch_col_wide(pl, nt, grid, block, lds, s, a)) return hipErrorInvalidValue;
        return hipGetLastError();
    }
    const bool ok = launch_col_part_a(pl, grid, block, lds, s, a) ||
           launch_col_part_b(pl, grid, block, lds, s, a) ||
           launch_col_part_c(pl, grid, block, lds, s, a);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_xpass_batch(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                              const int64_t* starts, int n) {
    const bool ok = launch_xb_part_a(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_xb_part_b(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_xb_part_c(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_xb_part_d(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_xb_part_e(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_xb_part_f(E, grid, block, lds, s, mode, ch, starts, n);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_colpass_batch(int E, int nt, bool mixed, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode,
                                const ColChunk* ch, const int64_t* starts, int n) {
    if (mixed) {
        if (nt != kMaxThreads || !launch_colb_mixed(E, grid, block, lds, s, mode, ch, starts, n))
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (nt != kMaxThreads) {
        if (!launch_colb_wide(E, nt, grid, block, lds, s, mode, ch, starts, n)) return hipErrorInvalidValue;
        return hipGetLastError();
    }
    const bool ok = launch_colb_part_a(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_colb_part_b(E, grid, block, lds, s, mode, ch, starts, n) ||
           launch_colb_part_c(E, grid, block, lds, s, mode, ch, starts, n);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

bool e_supported(int E) {
    switch (E) {
        case 1:
        case 2:
        case 3:
        case 4:
        case 5:
        case 6:
        case 8:
        case 9:
        case 10:
        case 12:
        case 15:
        case 16:
        case 18:
        case 20:
        case 24:
        case 25:
        case 27:
        case 30:
        case 32:
            return true;
        default:
            return false;
    }
}
}  // namespace pk

// pk_kern_col_mixed_b.hip -- strided-axis transform kernels for plans with
// stages fed from LDS, values per thread 12, 15, 16, 18, 20, 24, and the
// dispatch over these and pk_kern_col_mixed_a.hip.
#include "pk_launch.hpp"

namespace pk {
bool launch_col_mixed_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool launch_colb_mixed_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                         const int64_t* starts, int n);

bool col_mixed_available(int E, int nt) {
    if (nt != kMaxThreads) return false;
    switch (E) {
        case 4: case 5: case 6: case 8: case 9: case 10: case 12: case 15: case 16: case 18: case 20: case 24:
            return true;
        default:
            return false;
    }
}

bool launch_col_mixed(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
#define PK_MIXED_CASE(E) \
    case E: launch_modes_col<E, kMaxThreads, true>(grid, block, lds, s, a, pl); return true;
        PK_MIXED_CASE(12) PK_MIXED_CASE(15) PK_MIXED_CASE(16) PK_MIXED_CASE(18) PK_MIXED_CASE(20) PK_MIXED_CASE(24)
#undef PK_MIXED_CASE
        default: return launch_col_mixed_a(pl, grid, block, lds, s, a);
    }
}
bool launch_colb_mixed(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                       const int64_t* starts, int n) {
    switch (E) {
#define PK_MIXED_CASE(E) \
    case E: launch_modes_colb<E, kMaxThreads, true>(grid, block, lds, s, mode, ch, starts, n); return true;
        PK_MIXED_CASE(12) PK_MIXED_CASE(15) PK_MIXED_CASE(16) PK_MIXED_CASE(18) PK_MIXED_CASE(20) PK_MIXED_CASE(24)
#undef PK_MIXED_CASE
        default: return launch_colb_mixed_a(E, grid, block, lds, s, mode, ch, starts, n);
    }
}
}  // namespace pk

// pk_kern_col_mixed_a.hip -- strided-axis transform kernels (single-volume and
// batched, 256-thread blocks) for plans with stages fed from LDS, values per
// thread 4, 5, 6, 8, 9, 10; pk_kern_col_mixed_b.hip holds 12-24 and the
// dispatch over both.
#include "pk_launch.hpp"

namespace pk {
bool launch_col_mixed_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
#define PK_MIXED_CASE(E) \
    case E: launch_modes_col<E, kMaxThreads, true>(grid, block, lds, s, a, pl); return true;
        PK_MIXED_CASE(4) PK_MIXED_CASE(5) PK_MIXED_CASE(6) PK_MIXED_CASE(8) PK_MIXED_CASE(9) PK_MIXED_CASE(10)
#undef PK_MIXED_CASE
        default: return false;
    }
}
bool launch_colb_mixed_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                         const int64_t* starts, int n) {
    switch (E) {
#define PK_MIXED_CASE(E) \
    case E: launch_modes_colb<E, kMaxThreads, true>(grid, block, lds, s, mode, ch, starts, n); return true;
        PK_MIXED_CASE(4) PK_MIXED_CASE(5) PK_MIXED_CASE(6) PK_MIXED_CASE(8) PK_MIXED_CASE(9) PK_MIXED_CASE(10)
#undef PK_MIXED_CASE
        default: return false;
    }
}
}  // namespace pk

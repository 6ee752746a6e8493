// pk_kern_xrow_02.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 81/9, 90/10, 96/12, 100/10, 108/12, 120/8.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 81 && E == 9) {
        launch_modes_xrow<81, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 90 && E == 10) {
        launch_modes_xrow<90, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 96 && E == 12) {
        launch_modes_xrow<96, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 100 && E == 10) {
        launch_modes_xrow<100, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 108 && E == 12) {
        launch_modes_xrow<108, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 120 && E == 8) {
        launch_modes_xrow<120, 8>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 81 && E == 9) {
        launch_modes_xrowb<81, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 90 && E == 10) {
        launch_modes_xrowb<90, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 96 && E == 12) {
        launch_modes_xrowb<96, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 100 && E == 10) {
        launch_modes_xrowb<100, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 108 && E == 12) {
        launch_modes_xrowb<108, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 120 && E == 8) {
        launch_modes_xrowb<120, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_02(int L, int E) { return (L == 81 && E == 9) || (L == 90 && E == 10) || (L == 96 && E == 12) || (L == 100 && E == 10) || (L == 108 && E == 12) || (L == 120 && E == 8); }
}  // namespace pk

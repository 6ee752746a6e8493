// pk_kern_xrow_01.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 72/18, 80/10, 96/12, 120/30, 128/8, 144/18.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 72 && E == 18) {
        launch_modes_xrow<72, 18>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 80 && E == 10) {
        launch_modes_xrow<80, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 96 && E == 12) {
        launch_modes_xrow<96, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 120 && E == 30) {
        launch_modes_xrow<120, 30>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrow<128, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 144 && E == 18) {
        launch_modes_xrow<144, 18>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 72 && E == 18) {
        launch_modes_xrowb<72, 18>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 80 && E == 10) {
        launch_modes_xrowb<80, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 96 && E == 12) {
        launch_modes_xrowb<96, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 120 && E == 30) {
        launch_modes_xrowb<120, 30>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrowb<128, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 144 && E == 18) {
        launch_modes_xrowb<144, 18>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_01(int L, int E) { return (L == 72 && E == 18) || (L == 80 && E == 10) || (L == 96 && E == 12) || (L == 120 && E == 30) || (L == 128 && E == 8) || (L == 144 && E == 18); }
}  // namespace pk

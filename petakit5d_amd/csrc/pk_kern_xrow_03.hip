// pk_kern_xrow_03.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 125/5, 128/8, 135/9, 144/9, 150/10, 160/10.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_03(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 125 && E == 5) {
        launch_modes_xrow<125, 5>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrow<128, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 135 && E == 9) {
        launch_modes_xrow<135, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 144 && E == 9) {
        launch_modes_xrow<144, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 150 && E == 10) {
        launch_modes_xrow<150, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 160 && E == 10) {
        launch_modes_xrow<160, 10>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_03(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 125 && E == 5) {
        launch_modes_xrowb<125, 5>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrowb<128, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 135 && E == 9) {
        launch_modes_xrowb<135, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 144 && E == 9) {
        launch_modes_xrowb<144, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 150 && E == 10) {
        launch_modes_xrowb<150, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 160 && E == 10) {
        launch_modes_xrowb<160, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_03(int L, int E) { return (L == 125 && E == 5) || (L == 128 && E == 8) || (L == 135 && E == 9) || (L == 144 && E == 9) || (L == 150 && E == 10) || (L == 160 && E == 10); }
}  // namespace pk

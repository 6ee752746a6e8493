// pk_kern_xrow_05.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 240/8, 243/9, 250/10, 256/8, 270/9, 288/9.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_05(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 240 && E == 8) {
        launch_modes_xrow<240, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 243 && E == 9) {
        launch_modes_xrow<243, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 250 && E == 10) {
        launch_modes_xrow<250, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrow<256, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 270 && E == 9) {
        launch_modes_xrow<270, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 288 && E == 9) {
        launch_modes_xrow<288, 9>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_05(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 240 && E == 8) {
        launch_modes_xrowb<240, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 243 && E == 9) {
        launch_modes_xrowb<243, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 250 && E == 10) {
        launch_modes_xrowb<250, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrowb<256, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 270 && E == 9) {
        launch_modes_xrowb<270, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 288 && E == 9) {
        launch_modes_xrowb<288, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_05(int L, int E) { return (L == 240 && E == 8) || (L == 243 && E == 9) || (L == 250 && E == 10) || (L == 256 && E == 8) || (L == 270 && E == 9) || (L == 288 && E == 9); }
}  // namespace pk

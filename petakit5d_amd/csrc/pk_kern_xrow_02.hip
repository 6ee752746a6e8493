// pk_kern_xrow_02.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 160/10, 192/12, 240/30, 256/8, 288/18, 320/10.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 160 && E == 10) {
        launch_modes_xrow<160, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 192 && E == 12) {
        launch_modes_xrow<192, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 240 && E == 30) {
        launch_modes_xrow<240, 30>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrow<256, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 288 && E == 18) {
        launch_modes_xrow<288, 18>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 320 && E == 10) {
        launch_modes_xrow<320, 10>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 160 && E == 10) {
        launch_modes_xrowb<160, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 192 && E == 12) {
        launch_modes_xrowb<192, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 240 && E == 30) {
        launch_modes_xrowb<240, 30>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrowb<256, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 288 && E == 18) {
        launch_modes_xrowb<288, 18>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 320 && E == 10) {
        launch_modes_xrowb<320, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_02(int L, int E) { return (L == 160 && E == 10) || (L == 192 && E == 12) || (L == 240 && E == 30) || (L == 256 && E == 8) || (L == 288 && E == 18) || (L == 320 && E == 10); }
}  // namespace pk

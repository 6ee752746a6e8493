// pk_kern_xrow_06.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 300/10, 320/10, 324/6, 360/12, 375/15, 384/12.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_06(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 300 && E == 10) {
        launch_modes_xrow<300, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 320 && E == 10) {
        launch_modes_xrow<320, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 324 && E == 6) {
        launch_modes_xrow<324, 6>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 360 && E == 12) {
        launch_modes_xrow<360, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 375 && E == 15) {
        launch_modes_xrow<375, 15>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 384 && E == 12) {
        launch_modes_xrow<384, 12>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_06(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 300 && E == 10) {
        launch_modes_xrowb<300, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 320 && E == 10) {
        launch_modes_xrowb<320, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 324 && E == 6) {
        launch_modes_xrowb<324, 6>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 360 && E == 12) {
        launch_modes_xrowb<360, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 375 && E == 15) {
        launch_modes_xrowb<375, 15>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 384 && E == 12) {
        launch_modes_xrowb<384, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_06(int L, int E) { return (L == 300 && E == 10) || (L == 320 && E == 10) || (L == 324 && E == 6) || (L == 360 && E == 12) || (L == 375 && E == 15) || (L == 384 && E == 12); }
}  // namespace pk

// pk_kern_xrow_07.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 400/8, 405/9, 432/8, 450/15, 480/8, 486/9.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_07(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 400 && E == 8) {
        launch_modes_xrow<400, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 405 && E == 9) {
        launch_modes_xrow<405, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 432 && E == 8) {
        launch_modes_xrow<432, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 450 && E == 15) {
        launch_modes_xrow<450, 15>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 480 && E == 8) {
        launch_modes_xrow<480, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 486 && E == 9) {
        launch_modes_xrow<486, 9>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_07(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 400 && E == 8) {
        launch_modes_xrowb<400, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 405 && E == 9) {
        launch_modes_xrowb<405, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 432 && E == 8) {
        launch_modes_xrowb<432, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 450 && E == 15) {
        launch_modes_xrowb<450, 15>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 480 && E == 8) {
        launch_modes_xrowb<480, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 486 && E == 9) {
        launch_modes_xrowb<486, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_07(int L, int E) { return (L == 400 && E == 8) || (L == 405 && E == 9) || (L == 432 && E == 8) || (L == 450 && E == 15) || (L == 480 && E == 8) || (L == 486 && E == 9); }
}  // namespace pk

// pk_kern_xrow_11.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) -135/9, -225/15, -243/9, -375/15, -405/9.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_11(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == -135 && E == 9) {
        launch_modes_xrow<135, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -225 && E == 15) {
        launch_modes_xrow<225, 15, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -243 && E == 9) {
        launch_modes_xrow<243, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -375 && E == 15) {
        launch_modes_xrow<375, 15, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -405 && E == 9) {
        launch_modes_xrow<405, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_11(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == -135 && E == 9) {
        launch_modes_xrowb<135, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -225 && E == 15) {
        launch_modes_xrowb<225, 15, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -243 && E == 9) {
        launch_modes_xrowb<243, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -375 && E == 15) {
        launch_modes_xrowb<375, 15, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -405 && E == 9) {
        launch_modes_xrowb<405, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_11(int L, int E) { return (L == -135 && E == 9) || (L == -225 && E == 15) || (L == -243 && E == 9) || (L == -375 && E == 15) || (L == -405 && E == 9); }
}  // namespace pk

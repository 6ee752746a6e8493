// pk_kern_xrow_10.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) -25/5, -27/9, -45/9, -75/15, -81/9, -125/5.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_10(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == -25 && E == 5) {
        launch_modes_xrow<25, 5, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -27 && E == 9) {
        launch_modes_xrow<27, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -45 && E == 9) {
        launch_modes_xrow<45, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -75 && E == 15) {
        launch_modes_xrow<75, 15, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -81 && E == 9) {
        launch_modes_xrow<81, 9, true>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -125 && E == 5) {
        launch_modes_xrow<125, 5, true>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_10(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == -25 && E == 5) {
        launch_modes_xrowb<25, 5, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -27 && E == 9) {
        launch_modes_xrowb<27, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -45 && E == 9) {
        launch_modes_xrowb<45, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -75 && E == 15) {
        launch_modes_xrowb<75, 15, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -81 && E == 9) {
        launch_modes_xrowb<81, 9, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -125 && E == 5) {
        launch_modes_xrowb<125, 5, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_10(int L, int E) { return (L == -25 && E == 5) || (L == -27 && E == 9) || (L == -45 && E == 9) || (L == -75 && E == 15) || (L == -81 && E == 9) || (L == -125 && E == 5); }
}  // namespace pk

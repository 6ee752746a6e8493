// pk_kern_xrow_01.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 54/6, 60/5, 64/8, 72/8, 75/15, 80/10.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 54 && E == 6) {
        launch_modes_xrow<54, 6>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 60 && E == 5) {
        launch_modes_xrow<60, 5>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrow<64, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 72 && E == 8) {
        launch_modes_xrow<72, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 75 && E == 15) {
        launch_modes_xrow<75, 15>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 80 && E == 10) {
        launch_modes_xrow<80, 10>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 54 && E == 6) {
        launch_modes_xrowb<54, 6>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 60 && E == 5) {
        launch_modes_xrowb<60, 5>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrowb<64, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 72 && E == 8) {
        launch_modes_xrowb<72, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 75 && E == 15) {
        launch_modes_xrowb<75, 15>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 80 && E == 10) {
        launch_modes_xrowb<80, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_01(int L, int E) { return (L == 54 && E == 6) || (L == 60 && E == 5) || (L == 64 && E == 8) || (L == 72 && E == 8) || (L == 75 && E == 15) || (L == 80 && E == 10); }
}  // namespace pk

// pk_kern_xrow_09.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 1024/8, 256/4, 256/16, 128/4, 512/16, -15/15.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_09(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 1024 && E == 8) {
        launch_modes_xrow<1024, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 256 && E == 4) {
        launch_modes_xrow<256, 4>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 256 && E == 16) {
        launch_modes_xrow<256, 16>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 128 && E == 4) {
        launch_modes_xrow<128, 4>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 512 && E == 16) {
        launch_modes_xrow<512, 16>(grid, lds, s, a, tw);
        return true;
    }
    if (L == -15 && E == 15) {
        launch_modes_xrow<15, 15, true>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_09(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 1024 && E == 8) {
        launch_modes_xrowb<1024, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 256 && E == 4) {
        launch_modes_xrowb<256, 4>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 256 && E == 16) {
        launch_modes_xrowb<256, 16>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 128 && E == 4) {
        launch_modes_xrowb<128, 4>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 512 && E == 16) {
        launch_modes_xrowb<512, 16>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == -15 && E == 15) {
        launch_modes_xrowb<15, 15, true>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_09(int L, int E) { return (L == 1024 && E == 8) || (L == 256 && E == 4) || (L == 256 && E == 16) || (L == 128 && E == 4) || (L == 512 && E == 16) || (L == -15 && E == 15); }
}  // namespace pk

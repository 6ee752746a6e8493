// pk_kern_xrow_08.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 500/10, 512/8, 576/18, 640/10, 768/12, 960/30.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_08(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 500 && E == 10) {
        launch_modes_xrow<500, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 512 && E == 8) {
        launch_modes_xrow<512, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 576 && E == 18) {
        launch_modes_xrow<576, 18>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 640 && E == 10) {
        launch_modes_xrow<640, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 768 && E == 12) {
        launch_modes_xrow<768, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 960 && E == 30) {
        launch_modes_xrow<960, 30>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_08(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 500 && E == 10) {
        launch_modes_xrowb<500, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 512 && E == 8) {
        launch_modes_xrowb<512, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 576 && E == 18) {
        launch_modes_xrowb<576, 18>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 640 && E == 10) {
        launch_modes_xrowb<640, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 768 && E == 12) {
        launch_modes_xrowb<768, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 960 && E == 30) {
        launch_modes_xrowb<960, 30>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_08(int L, int E) { return (L == 500 && E == 10) || (L == 512 && E == 8) || (L == 576 && E == 18) || (L == 640 && E == 10) || (L == 768 && E == 12) || (L == 960 && E == 30); }
}  // namespace pk

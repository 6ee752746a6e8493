// pk_kern_xrow_00.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 32/8, 36/18, 40/10, 48/12, 60/30, 64/8.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_00(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 32 && E == 8) {
        launch_modes_xrow<32, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 36 && E == 18) {
        launch_modes_xrow<36, 18>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 40 && E == 10) {
        launch_modes_xrow<40, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 48 && E == 12) {
        launch_modes_xrow<48, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 60 && E == 30) {
        launch_modes_xrow<60, 30>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrow<64, 8>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_00(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 32 && E == 8) {
        launch_modes_xrowb<32, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 36 && E == 18) {
        launch_modes_xrowb<36, 18>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 40 && E == 10) {
        launch_modes_xrowb<40, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 48 && E == 12) {
        launch_modes_xrowb<48, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 60 && E == 30) {
        launch_modes_xrowb<60, 30>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrowb<64, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_00(int L, int E) { return (L == 32 && E == 8) || (L == 36 && E == 18) || (L == 40 && E == 10) || (L == 48 && E == 12) || (L == 60 && E == 30) || (L == 64 && E == 8); }
}  // namespace pk

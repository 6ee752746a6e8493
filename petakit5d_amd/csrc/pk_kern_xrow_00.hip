// pk_kern_xrow_00.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 32/8, 36/6, 40/10, 45/9, 48/12, 50/10.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_00(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 32 && E == 8) {
        launch_modes_xrow<32, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 36 && E == 6) {
        launch_modes_xrow<36, 6>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 40 && E == 10) {
        launch_modes_xrow<40, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 45 && E == 9) {
        launch_modes_xrow<45, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 48 && E == 12) {
        launch_modes_xrow<48, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 50 && E == 10) {
        launch_modes_xrow<50, 10>(grid, lds, s, a, tw);
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
    if (L == 36 && E == 6) {
        launch_modes_xrowb<36, 6>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 40 && E == 10) {
        launch_modes_xrowb<40, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 45 && E == 9) {
        launch_modes_xrowb<45, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 48 && E == 12) {
        launch_modes_xrowb<48, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 50 && E == 10) {
        launch_modes_xrowb<50, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_00(int L, int E) { return (L == 32 && E == 8) || (L == 36 && E == 6) || (L == 40 && E == 10) || (L == 45 && E == 9) || (L == 48 && E == 12) || (L == 50 && E == 10); }
}  // namespace pk

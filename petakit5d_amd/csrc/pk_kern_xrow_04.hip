// pk_kern_xrow_04.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) 162/9, 180/9, 192/12, 200/10, 216/8, 225/15.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_04(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 162 && E == 9) {
        launch_modes_xrow<162, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 180 && E == 9) {
        launch_modes_xrow<180, 9>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 192 && E == 12) {
        launch_modes_xrow<192, 12>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 200 && E == 10) {
        launch_modes_xrow<200, 10>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 216 && E == 8) {
        launch_modes_xrow<216, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 225 && E == 15) {
        launch_modes_xrow<225, 15>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_04(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {
    if (L == 162 && E == 9) {
        launch_modes_xrowb<162, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 180 && E == 9) {
        launch_modes_xrowb<180, 9>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 192 && E == 12) {
        launch_modes_xrowb<192, 12>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 200 && E == 10) {
        launch_modes_xrowb<200, 10>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 216 && E == 8) {
        launch_modes_xrowb<216, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 225 && E == 15) {
        launch_modes_xrowb<225, 15>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_04(int L, int E) { return (L == 162 && E == 9) || (L == 180 && E == 9) || (L == 192 && E == 12) || (L == 200 && E == 10) || (L == 216 && E == 8) || (L == 225 && E == 15); }
}  // namespace pk

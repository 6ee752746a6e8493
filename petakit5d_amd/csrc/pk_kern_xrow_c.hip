// pk_kern_xrow_c.hip -- compile-time row-transform kernels (other values-per-thread counts, selectable with PK_FFT_EX for launch-shape sweeps), single-volume
// and batched launches.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_c(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
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
    return false;
}
bool launch_xrowb_c(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                     const int64_t* starts, int n) {
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
    return false;
}
bool xrow_available_c(int L, int E) { return (L == 256 && E == 4) || (L == 256 && E == 16) || (L == 128 && E == 4) || (L == 512 && E == 16); }
}  // namespace pk

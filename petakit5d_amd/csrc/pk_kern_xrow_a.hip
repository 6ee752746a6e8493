// pk_kern_xrow_a.hip -- compile-time row-transform kernels (row lengths n1 = 64 to 512, 8 values per thread), single-volume
// and batched launches.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_a(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 32 && E == 8) {
        launch_modes_xrow<32, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrow<64, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrow<128, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrow<256, 8>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_a(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                     const int64_t* starts, int n) {
    if (L == 32 && E == 8) {
        launch_modes_xrowb<32, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 64 && E == 8) {
        launch_modes_xrowb<64, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 128 && E == 8) {
        launch_modes_xrowb<128, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 256 && E == 8) {
        launch_modes_xrowb<256, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_a(int L, int E) { return (L == 32 && E == 8) || (L == 64 && E == 8) || (L == 128 && E == 8) || (L == 256 && E == 8); }
}  // namespace pk

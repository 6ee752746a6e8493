// pk_kern_xrow_b.hip -- compile-time row-transform kernels (row lengths n1 = 1024 and 2048, 8 values per thread), single-volume
// and batched launches.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_b(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    if (L == 512 && E == 8) {
        launch_modes_xrow<512, 8>(grid, lds, s, a, tw);
        return true;
    }
    if (L == 1024 && E == 8) {
        launch_modes_xrow<1024, 8>(grid, lds, s, a, tw);
        return true;
    }
    return false;
}
bool launch_xrowb_b(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                     const int64_t* starts, int n) {
    if (L == 512 && E == 8) {
        launch_modes_xrowb<512, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    if (L == 1024 && E == 8) {
        launch_modes_xrowb<1024, 8>(grid, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
bool xrow_available_b(int L, int E) { return (L == 512 && E == 8) || (L == 1024 && E == 8); }
}  // namespace pk

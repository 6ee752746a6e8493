// pk_kern_x_d.hip -- instantiations of the dim-1 (row) transform kernels (single-volume
// and batched) for values-per-thread counts 20, 24; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_d(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 20:
            launch_modes_x<20>(grid, block, lds, s, a, pl);
            break;
        case 24:
            launch_modes_x<24>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_xb_part_d(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 20:
            launch_modes_xb<20>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 24:
            launch_modes_xb<24>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

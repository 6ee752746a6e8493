// pk_kern_x_b.hip -- instantiations of the dim-1 (row) transform kernels (single-volume
// and batched) for values-per-thread counts 9, 10, 12, 15; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_b(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 9:
            launch_modes_x<9>(grid, block, lds, s, a, pl);
            break;
        case 10:
            launch_modes_x<10>(grid, block, lds, s, a, pl);
            break;
        case 12:
            launch_modes_x<12>(grid, block, lds, s, a, pl);
            break;
        case 15:
            launch_modes_x<15>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_xb_part_b(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 9:
            launch_modes_xb<9>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 10:
            launch_modes_xb<10>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 12:
            launch_modes_xb<12>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 15:
            launch_modes_xb<15>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

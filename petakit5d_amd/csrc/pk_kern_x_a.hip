// pk_kern_x_a.hip -- instantiations of the dim-1 (row) transform kernels (single-volume
// and batched) for values-per-thread counts 1, 2, 3, 4, 5, 6, 8; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 1:
            launch_modes_x<1>(grid, block, lds, s, a, pl);
            break;
        case 2:
            launch_modes_x<2>(grid, block, lds, s, a, pl);
            break;
        case 3:
            launch_modes_x<3>(grid, block, lds, s, a, pl);
            break;
        case 4:
            launch_modes_x<4>(grid, block, lds, s, a, pl);
            break;
        case 5:
            launch_modes_x<5>(grid, block, lds, s, a, pl);
            break;
        case 6:
            launch_modes_x<6>(grid, block, lds, s, a, pl);
            break;
        case 8:
            launch_modes_x<8>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_xb_part_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const XChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 1:
            launch_modes_xb<1>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 2:
            launch_modes_xb<2>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 3:
            launch_modes_xb<3>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 4:
            launch_modes_xb<4>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 5:
            launch_modes_xb<5>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 6:
            launch_modes_xb<6>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 8:
            launch_modes_xb<8>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

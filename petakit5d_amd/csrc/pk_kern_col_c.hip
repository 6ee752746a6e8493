// pk_kern_col_c.hip -- instantiations of the strided-axis (dims 2, 3) transform kernels (single-volume
// and batched) for values-per-thread counts 24, 25, 27, 30, 32; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_col_part_c(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
        case 24:
            launch_modes_col<24>(grid, block, lds, s, a, pl);
            break;
        case 25:
            launch_modes_col<25>(grid, block, lds, s, a, pl);
            break;
        case 27:
            launch_modes_col<27>(grid, block, lds, s, a, pl);
            break;
        case 30:
            launch_modes_col<30>(grid, block, lds, s, a, pl);
            break;
        case 32:
            launch_modes_col<32>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_colb_part_c(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 24:
            launch_modes_colb<24>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 25:
            launch_modes_colb<25>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 27:
            launch_modes_colb<27>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 30:
            launch_modes_colb<30>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 32:
            launch_modes_colb<32>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

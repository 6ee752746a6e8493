// pk_kern_col_b.hip -- instantiations of the strided-axis (dims 2, 3) transform kernels (single-volume
// and batched) for values-per-thread counts 15, 16, 18, 20; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_col_part_b(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
        case 15:
            launch_modes_col<15>(grid, block, lds, s, a, pl);
            break;
        case 16:
            launch_modes_col<16>(grid, block, lds, s, a, pl);
            break;
        case 18:
            launch_modes_col<18>(grid, block, lds, s, a, pl);
            break;
        case 20:
            launch_modes_col<20>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_colb_part_b(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 15:
            launch_modes_colb<15>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 16:
            launch_modes_colb<16>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 18:
            launch_modes_colb<18>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 20:
            launch_modes_colb<20>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

// pk_kern_col_a.hip -- instantiations of the strided-axis (dims 2, 3) transform kernels (single-volume
// and batched) for values-per-thread counts 1, 2, 3, 4, 5, 6, 8, 9, 10, 12; the instantiations are split
// over several translation units so they compile in parallel.
#include "pk_launch.hpp"

namespace pk {
bool launch_col_part_a(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
        case 1:
            launch_modes_col<1>(grid, block, lds, s, a, pl);
            break;
        case 2:
            launch_modes_col<2>(grid, block, lds, s, a, pl);
            break;
        case 3:
            launch_modes_col<3>(grid, block, lds, s, a, pl);
            break;
        case 4:
            launch_modes_col<4>(grid, block, lds, s, a, pl);
            break;
        case 5:
            launch_modes_col<5>(grid, block, lds, s, a, pl);
            break;
        case 6:
            launch_modes_col<6>(grid, block, lds, s, a, pl);
            break;
        case 8:
            launch_modes_col<8>(grid, block, lds, s, a, pl);
            break;
        case 9:
            launch_modes_col<9>(grid, block, lds, s, a, pl);
            break;
        case 10:
            launch_modes_col<10>(grid, block, lds, s, a, pl);
            break;
        case 12:
            launch_modes_col<12>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}

bool launch_colb_part_a(int E, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
           const int64_t* starts, int n) {
    switch (E) {
        case 1:
            launch_modes_colb<1>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 2:
            launch_modes_colb<2>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 3:
            launch_modes_colb<3>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 4:
            launch_modes_colb<4>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 5:
            launch_modes_colb<5>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 6:
            launch_modes_colb<6>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 8:
            launch_modes_colb<8>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 9:
            launch_modes_colb<9>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 10:
            launch_modes_colb<10>(grid, block, lds, s, mode, ch, starts, n);
            break;
        case 12:
            launch_modes_colb<12>(grid, block, lds, s, mode, ch, starts, n);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

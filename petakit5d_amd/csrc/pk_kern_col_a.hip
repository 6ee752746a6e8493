// pk_kern_col_a.hip -- instantiations of the strided-axis (dims 2, 3) transform kernel for
// values-per-thread counts 1, 2, 3, 4, 5, 6, 8, 9, 10, 12 (generated split: the kernels compile in
// parallel translation units).
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
}  // namespace pk

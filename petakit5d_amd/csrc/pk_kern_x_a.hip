// pk_kern_x_a.hip -- instantiations of the dim-1 (row) transform kernel for
// values-per-thread counts 1, 2, 3, 4, 5, 6, 8 (generated split: the kernels compile in
// parallel translation units).
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
}  // namespace pk

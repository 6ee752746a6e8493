// pk_kern_x_e.hip -- instantiations of the dim-1 (row) transform kernel for
// values-per-thread counts 25, 27 (generated split: the kernels compile in
// parallel translation units).
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_e(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 25:
            launch_modes_x<25>(grid, block, lds, s, a, pl);
            break;
        case 27:
            launch_modes_x<27>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

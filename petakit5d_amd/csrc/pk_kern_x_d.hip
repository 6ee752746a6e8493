// pk_kern_x_d.hip -- instantiations of the dim-1 (row) transform kernel for
// values-per-thread counts 20, 24 (generated split: the kernels compile in
// parallel translation units).
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
}  // namespace pk

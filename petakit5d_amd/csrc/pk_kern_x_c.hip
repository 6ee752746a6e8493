// pk_kern_x_c.hip -- instantiations of the dim-1 (row) transform kernel for
// values-per-thread counts 16, 18 (generated split: the kernels compile in
// parallel translation units).
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_c(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 16:
            launch_modes_x<16>(grid, block, lds, s, a, pl);
            break;
        case 18:
            launch_modes_x<18>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

// pk_kern_x_f.hip -- instantiations of the dim-1 (row) transform kernel for
// values-per-thread counts 30, 32 (generated split: the kernels compile in
// parallel translation units).
#include "pk_launch.hpp"

namespace pk {
bool launch_x_part_f(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
        case 30:
            launch_modes_x<30>(grid, block, lds, s, a, pl);
            break;
        case 32:
            launch_modes_x<32>(grid, block, lds, s, a, pl);
            break;
        default:
            return false;
    }
    return true;
}
}  // namespace pk

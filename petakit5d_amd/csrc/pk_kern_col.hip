// pk_kern_col.hip -- instantiations of the strided-axis (dims 2, 3) transform kernel.
#include "pk_launch.hpp"

namespace pk {
hipError_t launch_colpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    switch (pl.E) {
#define PK_C(EE)                                                        \
    case EE:                                                            \
        hipLaunchKernelGGL(k_colpass<EE>, grid, block, lds, s, a, pl);  \
        break;
        PK_E_LIST(PK_C)
#undef PK_C
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
}  // namespace pk

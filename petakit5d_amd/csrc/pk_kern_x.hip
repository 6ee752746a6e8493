// pk_kern_x.hip -- instantiations of the dim-1 (row) transform kernel.
#include "pk_launch.hpp"

namespace pk {
hipError_t launch_xpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a) {
    switch (pl.E) {
#define PK_X(EE)                                                       \
    case EE:                                                           \
        hipLaunchKernelGGL(k_xpass<EE>, grid, block, lds, s, a, pl);   \
        break;
        PK_E_LIST(PK_X)
#undef PK_X
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool e_supported(int E) {
    switch (E) {
#define PK_S(EE) case EE:
        PK_E_LIST(PK_S)
#undef PK_S
        return true;
        default:
            return false;
    }
}
}  // namespace pk

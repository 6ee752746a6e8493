// pk_launch.hpp -- host-side launchers of the templated transform kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pk_fft_dev.hpp"

// Values-per-thread counts compiled into the transform kernels.
#define PK_E_LIST(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32)

namespace pk {
hipError_t launch_xpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const XArgs& a);
hipError_t launch_colpass(const FftPlan& pl, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a);
bool e_supported(int E);
}  // namespace pk

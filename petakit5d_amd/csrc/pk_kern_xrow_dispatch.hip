// pk_kern_xrow_dispatch.hip -- dispatch over the compile-time row kernels listed in
// pk_kern_xrow_NN.hip (71 (L, E) pairs).  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {
bool launch_xrow_00(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_00(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_00(int L, int E);
bool launch_xrow_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_01(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_01(int L, int E);
bool launch_xrow_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_02(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_02(int L, int E);
bool launch_xrow_03(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_03(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_03(int L, int E);
bool launch_xrow_04(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_04(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_04(int L, int E);
bool launch_xrow_05(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_05(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_05(int L, int E);
bool launch_xrow_06(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_06(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_06(int L, int E);
bool launch_xrow_07(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_07(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_07(int L, int E);
bool launch_xrow_08(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_08(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_08(int L, int E);
bool launch_xrow_09(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_09(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_09(int L, int E);
bool launch_xrow_10(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_10(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_10(int L, int E);
bool launch_xrow_11(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);
bool launch_xrowb_11(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n);
bool xrow_available_11(int L, int E);

bool launch_xrow(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {
    return launch_xrow_00(L, E, grid, lds, s, a, tw) ||
           launch_xrow_01(L, E, grid, lds, s, a, tw) ||
           launch_xrow_02(L, E, grid, lds, s, a, tw) ||
           launch_xrow_03(L, E, grid, lds, s, a, tw) ||
           launch_xrow_04(L, E, grid, lds, s, a, tw) ||
           launch_xrow_05(L, E, grid, lds, s, a, tw) ||
           launch_xrow_06(L, E, grid, lds, s, a, tw) ||
           launch_xrow_07(L, E, grid, lds, s, a, tw) ||
           launch_xrow_08(L, E, grid, lds, s, a, tw) ||
           launch_xrow_09(L, E, grid, lds, s, a, tw) ||
           launch_xrow_10(L, E, grid, lds, s, a, tw) ||
           launch_xrow_11(L, E, grid, lds, s, a, tw);
}
bool launch_xrow_batch(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                       const int64_t* starts, int n) {
    return launch_xrowb_00(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_01(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_02(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_03(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_04(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_05(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_06(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_07(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_08(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_09(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_10(L, E, grid, lds, s, mode, ch, starts, n) ||
           launch_xrowb_11(L, E, grid, lds, s, mode, ch, starts, n);
}
bool xrow_available(int L, int E) {
    return xrow_available_00(L, E) ||
           xrow_available_01(L, E) ||
           xrow_available_02(L, E) ||
           xrow_available_03(L, E) ||
           xrow_available_04(L, E) ||
           xrow_available_05(L, E) ||
           xrow_available_06(L, E) ||
           xrow_available_07(L, E) ||
           xrow_available_08(L, E) ||
           xrow_available_09(L, E) ||
           xrow_available_10(L, E) ||
           xrow_available_11(L, E);
}
}  // namespace pk

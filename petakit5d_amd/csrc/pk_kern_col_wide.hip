// pk_kern_col_wide.hip -- strided-axis transform kernels compiled for 512- and
// 1024-thread blocks: long columns (1024-2048 points) keep 8-16-column tiles
// (64-128-byte row segments) with up to 160 KB of LDS per workgroup.
#include "pk_launch.hpp"

namespace pk {
bool col_wide_available(int E, int nt) { return (E == 16 && nt == 1024) || (E == 8 && nt == 1024) || (E == 16 && nt == 512) || (E == 24 && nt == 512) || (E == 32 && nt == 512); }

bool launch_col_wide(const FftPlan& pl, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s, const ColArgs& a) {
    if (pl.E == 16 && nt == 1024) {
        launch_modes_col<16, 1024>(grid, block, lds, s, a, pl);
        return true;
    }
    if (pl.E == 8 && nt == 1024) {
        launch_modes_col<8, 1024>(grid, block, lds, s, a, pl);
        return true;
    }
    if (pl.E == 16 && nt == 512) {
        launch_modes_col<16, 512>(grid, block, lds, s, a, pl);
        return true;
    }
    if (pl.E == 24 && nt == 512) {
        launch_modes_col<24, 512>(grid, block, lds, s, a, pl);
        return true;
    }
    if (pl.E == 32 && nt == 512) {
        launch_modes_col<32, 512>(grid, block, lds, s, a, pl);
        return true;
    }
    return false;
}

bool launch_colb_wide(int E, int nt, dim3 grid, dim3 block, size_t lds, hipStream_t s, int mode, const ColChunk* ch,
                      const int64_t* starts, int n) {
    if (E == 16 && nt == 1024) {
        launch_modes_colb<16, 1024>(grid, block, lds, s, mode, ch, starts, n);
        return true;
    }
    if (E == 8 && nt == 1024) {
        launch_modes_colb<8, 1024>(grid, block, lds, s, mode, ch, starts, n);
        return true;
    }
    if (E == 16 && nt == 512) {
        launch_modes_colb<16, 512>(grid, block, lds, s, mode, ch, starts, n);
        return true;
    }
    if (E == 24 && nt == 512) {
        launch_modes_colb<24, 512>(grid, block, lds, s, mode, ch, starts, n);
        return true;
    }
    if (E == 32 && nt == 512) {
        launch_modes_colb<32, 512>(grid, block, lds, s, mode, ch, starts, n);
        return true;
    }
    return false;
}
}  // namespace pk

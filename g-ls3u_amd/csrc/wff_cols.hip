// wff_cols.hip -- launcher of the column WFF2 kernel.
#include "wff_kernels.h"

namespace glsu {

hipError_t launch_cols(const WffPlan& p, hipStream_t s) {
  const int waves = p.cols_waves;
  const int smem = plan_smem_bytes(p.plany, waves, true);
  switch (p.plany) {
#define BODY                                                                                    \
  {                                                                                             \
    const int per = items_per_block<P>(waves);                                                  \
    const int items = p.nwx * p.W;                                                              \
    const int grid = (items + per - 1) / per;                                                   \
    switch (waves) {                                                                            \
      case 1:                                                                                   \
        hipLaunchKernelGGL((wff_cols_kernel<P, 1>), dim3(grid), dim3(64), smem, s, p.d_A,      \
                           p.d_D, p.d_Hy, p.d_twy, p.d_thr, p.W, p.H, p.nwx, p.nwy,             \
                           p.plane_elems);                                                      \
        break;                                                                                  \
      case 2:                                                                                   \
        hipLaunchKernelGGL((wff_cols_kernel<P, 2>), dim3(grid), dim3(128), smem, s, p.d_A,     \
                           p.d_D, p.d_Hy, p.d_twy, p.d_thr, p.W, p.H, p.nwx, p.nwy,             \
                           p.plane_elems);                                                      \
        break;                                                                                  \
      default:                                                                                  \
        hipLaunchKernelGGL((wff_cols_kernel<P, 4>), dim3(grid), dim3(256), smem, s, p.d_A,     \
                           p.d_D, p.d_Hy, p.d_twy, p.d_thr, p.W, p.H, p.nwx, p.nwy,             \
                           p.plane_elems);                                                      \
        break;                                                                                  \
    }                                                                                           \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace glsu

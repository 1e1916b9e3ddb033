// wff_rows_fwd.hip -- launcher of the row-forward WFF2 kernel.
#include "wff_kernels.h"

namespace glsu {

hipError_t launch_rows_fwd(const WffPlan& p, const float2* f, hipStream_t s) {
  const int waves = p.rows_waves;
  const int smem = plan_smem_bytes(p.planx, waves);
  switch (p.planx) {
#define BODY                                                                                   \
  {                                                                                            \
    const int per = items_per_block<P>(waves);                                                 \
    const int grid = (p.H + per - 1) / per;                                                    \
    if (waves == 1)                                                                            \
      hipLaunchKernelGGL((wff_rows_fwd_kernel<P, 1>), dim3(grid), dim3(64), smem, s, f, p.d_A, \
                         p.d_Hx, p.d_twx, p.W, p.H, p.nwx, p.plane_elems);                     \
    else                                                                                       \
      hipLaunchKernelGGL((wff_rows_fwd_kernel<P, 2>), dim3(grid), dim3(128), smem, s, f,       \
                         p.d_A, p.d_Hx, p.d_twx, p.W, p.H, p.nwx, p.plane_elems);              \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace glsu

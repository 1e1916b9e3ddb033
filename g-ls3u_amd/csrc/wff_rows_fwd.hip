// wff_rows_fwd.hip -- launcher of the row-forward WFF2 kernel.
#include "wff_kernels.h"

namespace glsu {

hipError_t launch_rows_fwd(const WffPlan& p, const float2* f, hipStream_t s) {
  const int rw = p.rows_waves;
  const int smem = plan_smem_bytes(p.planx, rw);
  switch (p.planx) {
#define BODY                                                                              \
  if constexpr (P::T == 64) {                                                             \
    hipLaunchKernelGGL((wff_rows_fwd_kernel<P>), dim3(p.H), dim3(64 * rw), smem, s, f,    \
                       p.d_A, p.d_Hx, p.d_twx, p.W, p.H, p.nwx, p.plane_elems);           \
  } else {                                                                                \
    return hipErrorInvalidValue;                                                          \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace glsu

// wff_rows_inv.hip -- launchers of the row-inverse WFF2 kernel (filtered / DPRA update).
#include "wff_kernels.h"

namespace glsu {

template <int MODE>
static hipError_t launch_rows_inv_impl(const WffPlan& p, float2* out, const UpdateArgs& u,
                                       hipStream_t s, const int* skip) {
  const int rw = p.rows_waves;
  const int smem = plan_smem_bytes(p.planx, rw);
  switch (p.planx) {
#define BODY                                                                                 \
  if constexpr (P::T == 64) {                                                                \
    hipLaunchKernelGGL((wff_rows_inv_kernel<P, MODE>), dim3(p.H), dim3(64 * rw), smem, s,    \
                       p.d_D, p.d_Hx, p.d_twx, p.W, p.H, p.nwx, p.plane_elems, p.scale, out, \
                       u, skip);                                                             \
  } else {                                                                                   \
    return hipErrorInvalidValue;                                                             \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rows_inv_filtered(const WffPlan& p, float2* out, hipStream_t s) {
  UpdateArgs u{};
  return launch_rows_inv_impl<0>(p, out, u, s, nullptr);
}

hipError_t launch_rows_inv_update(const WffPlan& p, const UpdateArgs& u, hipStream_t s,
                                  const int* skip) {
  return launch_rows_inv_impl<1>(p, nullptr, u, s, skip);
}

hipError_t launch_rows_inv_batch(const BatchClass& c, const int* frame, hipStream_t s) {
  const int rw = c.rows_waves;
  const int smem = plan_smem_bytes(c.planx, rw);
  switch (c.planx) {
#define BODY                                                                                  \
  if constexpr (P::T == 64) {                                                                 \
    hipLaunchKernelGGL((wff_rows_inv_batch_kernel<P>), dim3(c.grid_rows), dim3(64 * rw), smem, \
                       s, c.tab, c.nseg, c.nrows, c.Hx, c.twx, c.nwx, c.scale, frame);        \
  } else {                                                                                    \
    return hipErrorInvalidValue;                                                              \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace glsu

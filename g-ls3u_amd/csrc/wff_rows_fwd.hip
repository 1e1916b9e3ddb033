// wff_rows_fwd.hip -- launcher of the row-forward WFF2 kernel.
#include "wff_kernels.h"

namespace glsu {

hipError_t launch_rows_fwd(const WffPlan& p, const float2* f, hipStream_t s, const int* skip,
                           int ld) {
  if (ld == 0) ld = p.W;
  if (ld < p.W) return hipErrorInvalidValue;
  const int rw = p.rows_waves;
  const int smem = plan_smem_bytes(p.planx, rw);
  switch (p.planx) {
#define BODY                                                                                \
  if constexpr (P::T == 64) {                                                               \
    if (live_regs_fit<P>(p.W))                                                              \
      hipLaunchKernelGGL((wff_rows_fwd_kernel<P, kLiveRegs<P>>), dim3(p.H), dim3(64 * rw),  \
                         smem, s, f, ld, p.d_A, p.d_Hx, p.d_twx, p.W, p.H, p.nwx,           \
                         p.plane_elems, skip);                                              \
    else                                                                                    \
      hipLaunchKernelGGL((wff_rows_fwd_kernel<P>), dim3(p.H), dim3(64 * rw), smem, s, f, ld, \
                         p.d_A, p.d_Hx, p.d_twx, p.W, p.H, p.nwx, p.plane_elems, skip);     \
  } else {                                                                                  \
    return hipErrorInvalidValue;                                                            \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Resident-workgroup capacity of a kernel on the current device, as a
// persistent grid: a multiple of 8 (one share per XCD), at most `items`.
template <class K>
static int persistent_grid(K kernel, int block, int smem, int items) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, smem) != hipSuccess)
    return -1;
  const int cap = std::max(8, (cus * std::max(per, 1)) / 8 * 8);
  return items <= cap ? items : cap;
}

int batch_grid_rows(int planx, int rows_waves, int nrows) {
  const int smem = plan_smem_bytes(planx, rows_waves);
  switch (planx) {
#define BODY                                                                              \
  if constexpr (P::T == 64) {                                                             \
    return persistent_grid(wff_rows_fwd_batch_kernel<P>, 64 * rows_waves, smem, nrows);   \
  } else {                                                                                \
    return -1;                                                                            \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return -1;
  }
  return -1;
}

hipError_t launch_rows_fwd_batch(const BatchClass& c, hipStream_t s) {
  const int rw = c.rows_waves;
  const int smem = plan_smem_bytes(c.planx, rw);
  switch (c.planx) {
#define BODY                                                                                 \
  if constexpr (P::T == 64) {                                                                \
    if (live_regs_fit<P>(c.maxW))                                                            \
      hipLaunchKernelGGL((wff_rows_fwd_batch_kernel<P, kLiveRegs<P>>), dim3(c.grid_rows),    \
                         dim3(64 * rw), smem, s, c.tab, c.nseg, c.nrows, c.Hx, c.twx, c.nwx); \
    else                                                                                     \
      hipLaunchKernelGGL((wff_rows_fwd_batch_kernel<P>), dim3(c.grid_rows), dim3(64 * rw),   \
                         smem, s, c.tab, c.nseg, c.nrows, c.Hx, c.twx, c.nwx);               \
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

}  // namespace glsu

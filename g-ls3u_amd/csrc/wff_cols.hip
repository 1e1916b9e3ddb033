// wff_cols.hip -- launcher of the column WFF2 kernel.
#include "wff_kernels.h"

namespace glsu {

template <class P, int WAVES, int EV>
static void cols_launch(const WffPlan& p, int grid, int smem, hipStream_t s, const int* skip) {
  hipLaunchKernelGGL((wff_cols_kernel<P, WAVES, EV>), dim3(grid), dim3(64 * WAVES), smem, s,
                     p.d_A, p.d_D, p.d_Hy, p.d_twy, p.d_thr, p.W, p.H, p.nwx, p.nwy,
                     p.plane_elems, skip, p.d_nskip);
}
template <class P, int WAVES>
static void cols_launch_live(const WffPlan& p, int grid, int smem, hipStream_t s,
                             const int* skip) {
  if (live_regs_fit<P>(p.H)) cols_launch<P, WAVES, kLiveRegs<P>>(p, grid, smem, s, skip);
  else cols_launch<P, WAVES, P::E>(p, grid, smem, s, skip);
}

hipError_t launch_cols(const WffPlan& p, hipStream_t s, const int* skip) {
  const int waves = p.cols_waves;
  const int smem = plan_smem_bytes(p.plany, waves, true);
  switch (p.plany) {
#define BODY                                                                 \
  {                                                                          \
    const int per = items_per_block<P>(waves);                               \
    const int grid = p.nwx * ((p.W + per - 1) / per);                        \
    switch (waves) {                                                         \
      case 1: cols_launch_live<P, 1>(p, grid, smem, s, skip); break;         \
      case 2: cols_launch_live<P, 2>(p, grid, smem, s, skip); break;         \
      default: cols_launch_live<P, 4>(p, grid, smem, s, skip); break;        \
    }                                                                        \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int cols_items_per_block(int plany, int waves) {
  for (auto& q : fft_plans())
    if (q.id == plany) return waves * 64 / q.T;
  return 0;
}

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

int batch_grid_cols(int plany, int waves, int nblk) {
  const int smem = plan_smem_bytes(plany, waves, true);
  switch (plany) {
#define BODY                                                                              \
  switch (waves) {                                                                        \
    case 1:                                                                               \
      return persistent_grid(wff_cols_batch_kernel<P, 1>, 64, smem, nblk);                \
    case 2:                                                                               \
      return persistent_grid(wff_cols_batch_kernel<P, 2>, 128, smem, nblk);               \
    default:                                                                              \
      return persistent_grid(wff_cols_batch_kernel<P, 4>, 256, smem, nblk);               \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return -1;
  }
  return -1;
}

template <class P, int WAVES, int EV>
static void cols_batch_launch(const BatchClass& c, int smem, hipStream_t s) {
  hipLaunchKernelGGL((wff_cols_batch_kernel<P, WAVES, EV>), dim3(c.grid_cols), dim3(64 * WAVES),
                     smem, s, c.tab, c.nseg, c.ncolblk, c.Hy, c.twy, c.nwx, c.nwy);
}
template <class P, int WAVES>
static void cols_batch_launch_live(const BatchClass& c, int smem, hipStream_t s) {
  if (live_regs_fit<P>(c.maxH)) cols_batch_launch<P, WAVES, kLiveRegs<P>>(c, smem, s);
  else cols_batch_launch<P, WAVES, P::E>(c, smem, s);
}

hipError_t launch_cols_batch(const BatchClass& c, hipStream_t s) {
  const int waves = c.cols_waves;
  const int smem = plan_smem_bytes(c.plany, waves, true);
  switch (c.plany) {
#define BODY                                                     \
  switch (waves) {                                               \
    case 1: cols_batch_launch_live<P, 1>(c, smem, s); break;     \
    case 2: cols_batch_launch_live<P, 2>(c, smem, s); break;     \
    default: cols_batch_launch_live<P, 4>(c, smem, s); break;    \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace glsu

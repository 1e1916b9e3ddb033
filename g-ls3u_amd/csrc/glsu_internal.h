// glsu_internal.h -- host-side declarations shared by the library's .hip files.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

namespace glsu {

// ---- FFT plans compiled into the library --------------------------------
// N = T * E, T lanes per transform, E elements per lane.
struct FftPlanInfo {
  int id;
  int T, E, N;
  const char* radices;  // cross-lane radices, for reports
  int exchange;         // 0 every stage through LDS, 1 first stage through registers
};
const std::vector<FftPlanInfo>& fft_plans();
// Plans with N >= need, sorted by N (then by fewest LDS stages).
std::vector<int> fft_candidates(int need);

// ---- WFF2 device workspace / plan ---------------------------------------
struct WffPlan {
  int W = 0, H = 0;
  int sx = 0, sy = 0;
  int nwx = 0, nwy = 0;
  double sigx = 0, sigy = 0, wxl = 0, wxh = 0, wxi = 0, wyl = 0, wyh = 0, wyi = 0;
  double thr_param = -1;     // < 0: 6*sqrt(mean|f|^2/3) from the first call
  bool thr_ready = false;
  float scale = 0;           // wxi*wyi/(4 pi^2)
  int planx = -1, plany = -1;  // FFT plan ids along x (rows) and y (columns)
  int Nx = 0, Ny = 0;
  int cols_waves = 4, rows_waves = 1;  // launch shapes
  int ls_tile = 1;                      // DPRA least-squares tile shape (launch_dpra_ls)
  // device buffers
  // window spectra / N, lane-major: [w][j][lane] = bin kbase(lane) + koff(j).
  // Real: the window is even, so its modulated spectrum has no imaginary part.
  float* d_Hx = nullptr;    // [nwx][Nx]
  float* d_Hy = nullptr;    // [nwy][Ny]
  float2* d_twx = nullptr;  // W_Nx^k
  float2* d_twy = nullptr;  // W_Ny^k
  float2* d_A = nullptr;    // [nwx] tiled planes: row-convolved field
  float2* d_D = nullptr;    // [nwx] tiled planes: column-filtered field
  float* d_thr = nullptr;   // threshold (device scalar)
  float* d_red = nullptr;   // reduction scratch
  unsigned int* d_nskip = nullptr;  // diagnostics: skipped column transforms (null = off)
  size_t plane_elems = 0;   // elements per tiled plane
};

// WFR mode: the eight result maps (device, row-major H x W floats), the
// per-wx ridge index planes and the curvature taps and constants.
struct WfrOut {
  float *wx, *wy, *phase, *phase_comp, *b, *r, *cxx, *cyy;
};
struct WfrConst {
  float inv_sx, inv_sy;  // 1 / (sum gy * sum k^2 gx), 1 / (sum gx * sum k^2 gy)
  float sig2x, sig2y;    // sigma^2
  float q4x, q4y;        // 1 / (4 pi sigma^2)
};
struct WfrBufs {
  uint16_t* d_idx = nullptr;  // [nwx] tiled planes: wy index of the best pair
  float* d_gx = nullptr;      // exp(-k^2 / 2 sigx^2), k = -sx..sx
  float* d_gy = nullptr;
  WfrConst c{};
};
hipError_t launch_wfr(const WffPlan& p, const WfrBufs& w, const WfrOut& z, hipStream_t s);
hipError_t launch_wfr_cols(const WffPlan& p, const WfrBufs& w, hipStream_t s);  // D, idx planes
hipError_t launch_wfr_post(const WffPlan& p, const WfrBufs& w, const WfrOut& z, hipStream_t s);
size_t wfr_curv_lds_bytes(int sx, int sy);

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2;
// remap so that each XCD walks a contiguous range of work items (lines shared
// by neighbouring items stay in one L2).
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int k = bid & 7, local = bid >> 3;
  return k * q + (k < r ? k : r) + local;
}

constexpr int kTileCols = 8;  // columns per tile in the intermediate planes
constexpr int kRowWavesMax = 4;  // row kernels: one wavefront per transform, <= 4 per row

int wff_plan_init(WffPlan& p, int W, int H, double sigx, double wxl, double wxh, double wxi,
                  double sigy, double wyl, double wyh, double wyi, double thr, int planx,
                  int plany, int cols_waves, int rows_waves, std::string& err,
                  bool buffers = true);  // false: spectra/twiddle tables only
void wff_plan_free(WffPlan& p);
// The scalar part of wff_plan_init (sizes, grid, plans, launch shapes) with
// no allocation of any kind; *err names the failure.
int wff_plan_geometry(WffPlan& p, int W, int H, double sigx, double wxl, double wxh, double wxi,
                      double sigy, double wyl, double wyh, double wyi, double thr, int planx,
                      int plany, int cols_waves, int rows_waves, const char** err);
// Default plan for a padded length >= need (row plans: t64_only).
int plan_default(int need, bool t64_only);
// wff_plan_init's tables (windows, spectra, twiddles; the threshold when
// thr_param >= 0) computed on the device into p.d_Hx, d_Hy, d_twx, d_twy,
// d_thr; gx, gy: device scratch of 2 sx + 1 and 2 sy + 1 doubles.
hipError_t launch_plan_tables(const WffPlan& p, double* gx, double* gy, hipStream_t s);

// Kernel launchers (all asynchronous on `s`).
hipError_t launch_thr(const WffPlan& p, const float2* f, hipStream_t s);
// skip (optional device flag): when *skip != 0 the kernel exits at once;
// ld: elements between rows of f (0: p.W), e.g. a window of a wider field
hipError_t launch_rows_fwd(const WffPlan& p, const float2* f, hipStream_t s,
                           const int* skip = nullptr, int ld = 0);
hipError_t launch_cols(const WffPlan& p, hipStream_t s, const int* skip = nullptr);
// mode 0: out = filtered complex field (H*W float2)
hipError_t launch_rows_inv_filtered(const WffPlan& p, float2* out, hipStream_t s);
// mode 1: DPRA phase update fused into the inverse row pass
struct UpdateArgs {
  const float* phiref;
  const float* dref_hi;
  const float* dref_lo;
  float* phicurr;
  float* dphi_hi;
  float* dphi_lo;
  float* dphi_out;   // may be null
  float2* filtered;  // may be null
};
hipError_t launch_rows_inv_update(const WffPlan& p, const UpdateArgs& u, hipStream_t s,
                                  const int* skip = nullptr);

// DPRA phase update of pixel id from the filtered phase t = angle(F)
// (DPRA/dpra_cpu.cpp:244-251): phicurr = wrap(phiref + t) into (-pi, pi]
// with 2 pi split in two floats; dphi = dphiref + t carried as an unevaluated
// float pair (TwoSum, then Fast2Sum renormalisation).
// phase_step: the arithmetic on values already loaded (phiref, dphiref hi/lo);
// phase_store writes its results.  Kernels that update several pixels per
// thread load all inputs first: the output stores could alias the inputs, so
// a load after a store is not moved above it, and on this target the wait for
// that load also waits for the stores before it.
struct PhaseOut {
  float pc, nh, nl;
};
__device__ __forceinline__ PhaseOut phase_step(float pr, float hi, float lo, float t) {
  float pc = pr + t;
  if (pc > 3.14159265f) pc = (pc - 6.28318548f) + 1.74845553e-7f;
  else if (pc <= -3.14159265f) pc = (pc + 6.28318548f) - 1.74845553e-7f;
  const float s = hi + t;
  const float bp = s - hi;
  const float err = (hi - (s - bp)) + (t - bp);
  const float lo2 = lo + err;
  const float nh = s + lo2;
  const float nl = lo2 - (nh - s);
  return PhaseOut{pc, nh, nl};
}
__device__ __forceinline__ void phase_store(const UpdateArgs& u, size_t id, const PhaseOut& o) {
  u.phicurr[id] = o.pc;
  u.dphi_hi[id] = o.nh;
  u.dphi_lo[id] = o.nl;
  if (u.dphi_out) u.dphi_out[id] = o.nh;
}
__device__ __forceinline__ void phase_update(const UpdateArgs& u, size_t id, float t) {
  phase_store(u, id, phase_step(u.phiref[id], u.dref_hi[id], u.dref_lo[id], t));
}
// learned path: update from the network output (t = angle(out)); residual
// of the network's phase against a WFF2 window on a 64 x 64 tile
// skip: device flag, the update is skipped when *skip != 0 (the frame fell back)
hipError_t launch_net_update(const float2* out, const UpdateArgs& u, size_t n, hipStream_t s,
                             const int* skip);
// flag = rms > tol, not_flag = !flag; stats[0] += 1 (frames), stats[1] += flag,
// stats[2] += rms, stats[3] = max(stats[3], rms), stats[4] = rms (single-thread update)
hipError_t launch_tile_residual(const float2* net, int W, int tx, int ty, const float2* fwin,
                                int Ww, int ox, int oy, float tol, int* flag, int* not_flag,
                                float* stats, hipStream_t s);

// ---- batched launches over many images --------------------------------
// One image of a batched launch (device-resident table entry).  row0 / col0 /
// tile0 place the image in the item space of its plan class's row launches,
// column launches (in column-kernel blocks) and of the least-squares launch.
// phi / dhi / dlo are double buffers whose reference/current roles follow the
// frame counter and rr (see swap_parity).
struct BatchSeg {
  int W, H, rr;
  int row0, col0, tile0;
  float inv_n;               // 1 / (W*H), for the threshold
  int pad;
  const void* frames;        // [T][H][W] frames of this image
  float* dphi_out;           // [T][H][W] cumulative phase out (may be null)
  float* phi[2];
  float* dhi[2];
  float* dlo[2];
  float2* z;
  float2* A;
  float2* D;
  size_t plane;
  float* thr;                // < 0 until computed from the first frame
  float* red;                // 256 partial sums
};

// Plan-class launch data of the batched WFF2 stages.
// Registers per lane of a transform of E points per lane that a kernel treats
// as live when the image extent fits in them (the rest hold only zero inputs
// and discarded outputs): E - 4 from 16 points per lane up.
constexpr int live_regs(int E) { return E >= 16 ? E - 4 : E; }

struct BatchClass {
  const BatchSeg* tab;       // device table of this class's images (contiguous)
  int nseg;
  int nrows;                 // sum of H
  int ncolblk;               // sum over images of nwx * ceil(W / items per block)
  int planx, plany, cols_waves, rows_waves;
  int nwx, nwy;
  int maxW, maxH;            // largest image of the class (live-register choice)
  float scale;
  const float *Hx, *Hy;
  const float2 *twx, *twy;
  int grid_rows, grid_cols;  // persistent grid sizes (multiples of 8)
};
hipError_t launch_rows_fwd_batch(const BatchClass& c, hipStream_t s);
hipError_t launch_cols_batch(const BatchClass& c, hipStream_t s);
hipError_t launch_rows_inv_batch(const BatchClass& c, const int* frame, hipStream_t s);
// items per column-kernel block for a plan and wave count
int cols_items_per_block(int plany, int waves);
// occupancy-limited resident workgroups of each batched kernel on the device
int batch_grid_rows(int planx, int rows_waves, int nrows);
int batch_grid_cols(int plany, int cols_waves, int nblk);
hipError_t launch_dpra_ls_batch(const BatchSeg* tab, int nseg, int ntiles, int grid, int dtype,
                                const int* frame, hipStream_t s);
int dpra_ls_tiles(int W, int H);
int batch_grid_ls(int ntiles);
hipError_t launch_thr_batch(const BatchSeg* tab, int nseg, hipStream_t s);
hipError_t launch_frame_advance(int* frame, hipStream_t s);

// DPRA least squares: z = exp(j*Delta) from frame + phiref
enum InDtype { IN_U8 = 0, IN_F32 = 1, IN_F64 = 2 };
// tile: least-squares tile shape id (1 = 64 x 8 ... ; ls_tile_valid)
hipError_t launch_dpra_ls(const void* frame, int dtype, const float* phiref, float2* z, int W,
                          int H, int pitch, hipStream_t s, int tile = 1);
bool ls_tile_valid(int id);
// synthetic fringe frames (device generator, same definition as the oracle's)
hipError_t launch_gen_frames(float* out, int W, int H, int t0, int T, uint64_t seed, double amp0,
                             double amp1, double carrier, double noise, hipStream_t s);

}  // namespace glsu

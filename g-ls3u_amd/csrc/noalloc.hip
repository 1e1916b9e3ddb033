// noalloc.hip -- WFF2 of one field with no allocation: every intermediate in
// one file-scope device array.
//
// glsu_wft2_create + glsu_wft2_compute allocate the plan's tables and the
// row / column planes with hipMalloc when the handle is made.  Callers that
// cannot allocate (code under HIP graph capture, tight serving loops) get
// glsu_wft2_compute_noalloc instead: the window spectra, twiddles, threshold
// and reduction scratch and both plane sets live in `scratch`, a static
// __device__ array of kScratchBytes (each device holds its own copy from the
// library's load until the process ends), and the tables are computed on the
// device every call.  Calls that overlap on one device share that array and
// race; the caller orders them (one stream, or events).
#include <mutex>

#include "../../include/glsu.h"
#include "wff_kernels.h"

namespace glsu {
void set_error(const char* msg);  // api.hip: glsu_last_error()'s message
}

// 96 MiB: a 1024 x 1024 field with up to 5 wx values (two planes of 5 x 8 MiB
// each plus under 1 MiB of tables) fits.
constexpr size_t kScratchBytes = (size_t)96 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using namespace glsu;

constexpr int kMaxDevices = 64;

int fail_noalloc(int code, const char* msg);

std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of this device's copy of `scratch`.  In host code the
// name `scratch` is a host-side stand-in, never a device address: kernels and
// copies get this pointer.
int scratch_ptr(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    return fail_noalloc(GLSU_ERR_DEVICE, "no current HIP device");
  if (dev >= kMaxDevices) return fail_noalloc(GLSU_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail_noalloc(GLSU_ERR_DEVICE, "hipGetSymbolAddress of the scratch array failed");
    g_ptr[dev] = (unsigned char*)p;
  }
  *out = g_ptr[dev];
  return GLSU_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int fail_noalloc(int code, const char* msg) {
  set_error(msg);
  return code;
}

}  // namespace

// One WFF-mode WFT2 (init + compute) of a height x width complex64 field, with
// the plans and launch shapes glsu_wft2_create picks by default; no handle,
// no allocation.  See include/glsu.h.
extern "C" int glsu_wft2_compute_noalloc(int width, int height, const glsu_wff_params* pr,
                                         const void* d_f, void* d_filtered, void* hip_stream) {
  if (!pr || !d_f || !d_filtered) return fail_noalloc(GLSU_ERR_ARG, "null argument");
  const int W = width, H = height;
  if (W < 1 || H < 1) return fail_noalloc(GLSU_ERR_ARG, "invalid size");
  if (!(pr->wxi > 0) || !(pr->wyi > 0) || !(pr->sigma_x > 0) || !(pr->sigma_y > 0) ||
      !(pr->wxh >= pr->wxl) || !(pr->wyh >= pr->wyl))
    return fail_noalloc(GLSU_ERR_ARG, "invalid WFF parameters");
  const int sx = (int)std::round(3 * pr->sigma_x), sy = (int)std::round(3 * pr->sigma_y);
  // the reference constructor's padding rule (WFT/WFT2_CPU.cpp:710-724)
  if (W + 2 * sx > 4096 || H + 2 * sy > 4096)
    return fail_noalloc(GLSU_ERR_PADDING, "FFT padding is out of range [4096]");
  const int planx = plan_default(W + sx, true), plany = plan_default(H + sy, false);
  if (planx < 0 || plany < 0)
    return fail_noalloc(GLSU_ERR_PADDING, "no compiled FFT length covers the size + half window");
  WffPlan p;
  const char* err = nullptr;
  if (wff_plan_geometry(p, W, H, pr->sigma_x, pr->wxl, pr->wxh, pr->wxi, pr->sigma_y, pr->wyl,
                        pr->wyh, pr->wyi, pr->thr, planx, plany, -1, -1, &err))
    return fail_noalloc(GLSU_ERR_ARG, err ? err : "invalid plan");
  // carve the scratch array: windows, spectra, twiddles, threshold, partial
  // sums, then the two plane sets
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const size_t o_gx = take(sizeof(double) * (2 * p.sx + 1));
  const size_t o_gy = take(sizeof(double) * (2 * p.sy + 1));
  const size_t o_hx = take(sizeof(float) * (size_t)p.nwx * p.Nx);
  const size_t o_hy = take(sizeof(float) * (size_t)p.nwy * p.Ny);
  const size_t o_tx = take(sizeof(float2) * (size_t)p.Nx);
  const size_t o_ty = take(sizeof(float2) * (size_t)p.Ny);
  const size_t o_thr = take(sizeof(float));
  const size_t o_red = take(sizeof(float) * 256);
  const size_t o_a = take(sizeof(float2) * p.plane_elems * p.nwx);
  const size_t o_d = take(sizeof(float2) * p.plane_elems * p.nwx);
  if (off > kScratchBytes)
    return fail_noalloc(GLSU_ERR_UNSUPPORTED,
                        "field too large for the no-allocation scratch (96 MiB: e.g. 1024x1024 "
                        "with up to 5 wx values)");
  unsigned char* base = nullptr;
  if (int rc = scratch_ptr(&base)) return rc;
  p.d_Hx = (float*)(base + o_hx);
  p.d_Hy = (float*)(base + o_hy);
  p.d_twx = (float2*)(base + o_tx);
  p.d_twy = (float2*)(base + o_ty);
  p.d_thr = (float*)(base + o_thr);
  p.d_red = (float*)(base + o_red);
  p.d_A = (float2*)(base + o_a);
  p.d_D = (float2*)(base + o_d);
  hipStream_t s = (hipStream_t)hip_stream;
  const float2* f = (const float2*)d_f;
  hipError_t e = launch_plan_tables(p, (double*)(base + o_gx), (double*)(base + o_gy), s);
  // thr < 0: computed from this field (nothing persists between calls, so
  // every call is the reference's first call, WFT/WFT2_CPU.cpp:1014-1015)
  if (e == hipSuccess && p.thr_param < 0) e = launch_thr(p, f, s);
  if (e == hipSuccess) e = launch_rows_fwd(p, f, s);
  if (e == hipSuccess) e = launch_cols(p, s);
  if (e == hipSuccess) e = launch_rows_inv_filtered(p, (float2*)d_filtered, s);
  if (e != hipSuccess) return fail_noalloc(GLSU_ERR_DEVICE, hipGetErrorString(e));
  return GLSU_OK;
}

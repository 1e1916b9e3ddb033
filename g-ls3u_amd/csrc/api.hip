// api.hip -- C ABI (include/glsu.h) over the gfx950 kernels.
//
// Host-side restatement of the reference's class logic:
//   WFT2_CUDAF ctor/operator()   WFT/WFT2_CUDAf.cu:849-1173  (WFF only)
//   DPRA_CUDAF ctor/dpra_per_frame/update_ref_phi  DPRA/dpra_cudaf.cu:96-290
//   DPRA_CPU::operator()         DPRA/dpra_cpu.cpp:58-89 (frame loop, i % rr)
// Differences by design: status codes instead of printing and returning
// (size mismatch, DPRA/dpra_cpu.cpp:135-139) or throwing int -1 (padding,
// WFT/WFT2_CPU.cpp:69-73); no host synchronisation per frame unless the
// caller asks for a time; WFF parameters exposed instead of hard-coded.
#include "../../include/glsu.h"
#include "glsu_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <tuple>
#include <vector>

using namespace glsu;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace glsu {
void set_error(const char* msg) { g_err = msg; }
int unet_forward_on(glsu_unet* h, const float2* x, float2* out, int H, int W, hipStream_t s);
}  // namespace glsu

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(GLSU_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

static int check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return fail(GLSU_ERR_DEVICE, "no HIP device visible");
  return GLSU_OK;
}

static int elem_size(int dtype) {
  switch (dtype) {
    case GLSU_U8: return 1;
    case GLSU_F32: return 4;
    case GLSU_F64: return 8;
    default: return 0;
  }
}

// NULL selects the default (null) stream, which the reference's GPU classes use
// (DPRA/dpra_cudaf.cu:207-284 launches on stream 0); any other value is used
// as given and never destroyed by the library.
static int make_stream(void* user, hipStream_t* s, bool* own) {
  *s = (hipStream_t)user;
  *own = false;
  return GLSU_OK;
}

extern "C" const char* glsu_last_error(void) { return g_err.c_str(); }
extern "C" const char* glsu_version(void) { return "glsu 0.1 (gfx950)"; }

extern "C" int glsu_device_count(int* count) {
  if (!count) return fail(GLSU_ERR_ARG, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return GLSU_OK;
}

extern "C" int glsu_wff_params_default(glsu_wff_params* p) {
  if (!p) return fail(GLSU_ERR_ARG, "null params");
  // WFT/WFT2_CPU.cpp:43-54
  p->sigma_x = 10.0;
  p->wxl = -2 - 3.0 / p->sigma_x;
  p->wxi = 1.0 / p->sigma_x;
  p->wxh = 2 + 3.0 / p->sigma_x;
  p->sigma_y = 10.0;
  p->wyl = -2 - 3.0 / p->sigma_y;
  p->wyi = 1.0 / p->sigma_y;
  p->wyh = 2 + 3.0 / p->sigma_y;
  p->thr = -1;
  return GLSU_OK;
}

extern "C" int glsu_wfr_params_default(glsu_wff_params* p) {
  if (!p) return fail(GLSU_ERR_ARG, "null params");
  // WFT/WFT2_CPU.cpp:56-67 (WFR-mode constructor defaults)
  p->sigma_x = 10.0; p->wxl = -2; p->wxi = 0.025; p->wxh = 2;
  p->sigma_y = 10.0; p->wyl = -2.0; p->wyi = 0.025; p->wyh = 2.0;
  p->thr = -1;
  return GLSU_OK;
}

extern "C" int glsu_wff_params_dpra_reference(glsu_wff_params* p) {
  if (!p) return fail(GLSU_ERR_ARG, "null params");
  // DPRA/dpra_cpu.cpp:23-25: float literals widened to double
  p->sigma_x = 20; p->wxl = (double)-0.15f; p->wxh = (double)0.15f; p->wxi = (double)0.05f;
  p->sigma_y = 20; p->wyl = (double)-0.15f; p->wyh = (double)0.15f; p->wyi = (double)0.05f;
  p->thr = 15;
  return GLSU_OK;
}

extern "C" int glsu_fft_plans(int* ids, int* lanes, int* elems, int* lengths, int cap, int* count) {
  const auto& v = fft_plans();
  int n = 0;
  for (auto& p : v) {
    if (n < cap) {
      if (ids) ids[n] = p.id;
      if (lanes) lanes[n] = p.T;
      if (elems) elems[n] = p.E;
      if (lengths) lengths[n] = p.N;
    }
    ++n;
  }
  if (count) *count = n;
  return GLSU_OK;
}

extern "C" int glsu_fft_plan_exchange(int id) {
  for (auto& p : fft_plans())
    if (p.id == id) return p.exchange;
  return fail(GLSU_ERR_ARG, "unknown FFT plan id %d", id);
}

// Padded size rule of the reference (WFT/WFT2_CPU.cpp:710-724 + WFT/WFT.h:18-32):
// the constructor fails when OPT(H + 2s) or OPT(W + 2s) exceeds 4096.  The
// library pads differently (any N >= size + s is exact) but keeps the same
// acceptance rule so that callers see the same failures.
static bool reference_padding_ok(int W, int H, int sx, int sy) {
  return W + 2 * sx <= 4096 && H + 2 * sy <= 4096;
}

// Row transforms (x axis) use one wavefront per transform (T == 64 plans);
// column transforms may pack several narrower transforms per wavefront.
static int default_plan(int need, bool t64_only) { return plan_default(need, t64_only); }

// Host-only validation (no device needed): sizes, parameters, padding rule.
static int validate(int W, int H, const glsu_wff_params* pr, int* planx, int* plany) {
  if (W < 1 || H < 1) return fail(GLSU_ERR_ARG, "invalid size %dx%d", W, H);
  if (!(pr->wxi > 0) || !(pr->wyi > 0) || !(pr->sigma_x > 0) || !(pr->sigma_y > 0) ||
      !(pr->wxh >= pr->wxl) || !(pr->wyh >= pr->wyl))
    return fail(GLSU_ERR_ARG, "invalid WFF parameters");
  const int sx = (int)std::round(3 * pr->sigma_x), sy = (int)std::round(3 * pr->sigma_y);
  if (!reference_padding_ok(W, H, sx, sy))
    return fail(GLSU_ERR_PADDING,
                "FFT padding is out of range [4096]. Shrink the size of either the image or the "
                "Gaussian Window!");
  if (*planx < 0) *planx = default_plan(W + sx, true);
  if (*plany < 0) *plany = default_plan(H + sy, false);
  if (*planx < 0 || *plany < 0)
    return fail(GLSU_ERR_PADDING, "no compiled FFT length covers %dx%d + half window", W, H);
  bool okx = false, oky = false;
  for (auto& q : fft_plans()) {
    if (q.id == *planx && q.N >= W + sx && q.T == 64) okx = true;
    if (q.id == *plany && q.N >= H + sy) oky = true;
  }
  if (!okx || !oky)
    return fail(GLSU_ERR_ARG,
                "FFT plan too short for image + half window (or a narrow plan on the x axis)");
  return GLSU_OK;
}

// Host-only validation of a tuning request (plan ids are checked by validate()).
static int validate_tuning(const glsu_tuning* t) {
  if (t && !(t->cols_waves < 0 || t->cols_waves == 1 || t->cols_waves == 2 || t->cols_waves == 4))
    return fail(GLSU_ERR_ARG, "cols_waves must be 1, 2 or 4 (got %d)", t->cols_waves);
  if (t && !(t->rows_waves < 0 || (t->rows_waves >= 1 && t->rows_waves <= kRowWavesMax)))
    return fail(GLSU_ERR_ARG, "rows_waves must be in 1..%d (got %d)", kRowWavesMax, t->rows_waves);
  if (t && t->ls_tile > 0 && !ls_tile_valid(t->ls_tile))
    return fail(GLSU_ERR_ARG, "unknown least-squares tile shape %d", t->ls_tile);
  return GLSU_OK;
}

static int plan_common(WffPlan& p, int W, int H, const glsu_wff_params* pr, const glsu_tuning* t) {
  int planx = t ? t->planx : -1, plany = t ? t->plany : -1;
  if (int rc = validate(W, H, pr, &planx, &plany)) return rc;
  std::string err;
  int rc = wff_plan_init(p, W, H, pr->sigma_x, pr->wxl, pr->wxh, pr->wxi, pr->sigma_y, pr->wyl,
                         pr->wyh, pr->wyi, pr->thr, planx, plany, t ? t->cols_waves : -1,
                         t ? t->rows_waves : -1, err);
  if (rc == -2) return fail(GLSU_ERR_ARG, "%s", err.c_str());
  if (rc) return fail(GLSU_ERR_DEVICE, "%s", err.c_str());
  if (t && t->ls_tile > 0) p.ls_tile = t->ls_tile;
  return GLSU_OK;
}

// WFR mode (WFT/WFT2_CPU.cpp:848-858): the constructor also fails when the
// curvature padding OPT(H + 3 sy) or OPT(W + 3 sx) exceeds 4096.  The library
// evaluates the curvatures directly, with LDS tiles that bound the window.
static int validate_wfr(int W, int H, const glsu_wff_params* pr) {
  const int sx = (int)std::round(3 * pr->sigma_x), sy = (int)std::round(3 * pr->sigma_y);
  if (W + 3 * sx > 4096 || H + 3 * sy > 4096)
    return fail(GLSU_ERR_PADDING,
                "FFT padding is out of range [4096]. Shrink the size of either the image or the "
                "Gaussian Window!");
  if (wfr_curv_lds_bytes(sx, sy) > 65536)
    return fail(GLSU_ERR_UNSUPPORTED, "WFR2 curvature window too large (sigma %g x %g)",
                pr->sigma_x, pr->sigma_y);
  const int nwy = (int)((pr->wyh - pr->wyl) * (1 / pr->wyi)) + 1;
  if (nwy >= 0xFFFF) return fail(GLSU_ERR_UNSUPPORTED, "WFR2 supports fewer than 65535 wy values");
  return GLSU_OK;
}

static int wfr_alloc(const WffPlan& p, WfrBufs& w) {
  HIP_TRY(hipMalloc((void**)&w.d_idx, sizeof(uint16_t) * p.plane_elems * p.nwx));
  std::vector<float> gx(2 * p.sx + 1), gy(2 * p.sy + 1);
  double sgx = 0, sgy = 0, sxx = 0, syy = 0;
  for (int k = -p.sx; k <= p.sx; ++k) {
    const double g = std::exp(-(double)k * k / 2 / p.sigx / p.sigx);
    gx[k + p.sx] = (float)g;
    sgx += g;
    sxx += (double)k * k * g;
  }
  for (int k = -p.sy; k <= p.sy; ++k) {
    const double g = std::exp(-(double)k * k / 2 / p.sigy / p.sigy);
    gy[k + p.sy] = (float)g;
    sgy += g;
    syy += (double)k * k * g;
  }
  HIP_TRY(hipMalloc((void**)&w.d_gx, sizeof(float) * gx.size()));
  HIP_TRY(hipMalloc((void**)&w.d_gy, sizeof(float) * gy.size()));
  HIP_TRY(hipMemcpy(w.d_gx, gx.data(), sizeof(float) * gx.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_gy, gy.data(), sizeof(float) * gy.size(), hipMemcpyHostToDevice));
  w.c.inv_sx = (float)(1.0 / (sgy * sxx));
  w.c.inv_sy = (float)(1.0 / (sgx * syy));
  w.c.sig2x = (float)(p.sigx * p.sigx);
  w.c.sig2y = (float)(p.sigy * p.sigy);
  w.c.q4x = (float)(0.25 / M_PI / (p.sigx * p.sigx));
  w.c.q4y = (float)(0.25 / M_PI / (p.sigy * p.sigy));
  return GLSU_OK;
}

static void wfr_free(WfrBufs& w) {
  hipFree(w.d_idx);
  hipFree(w.d_gx);
  hipFree(w.d_gy);
  w = WfrBufs{};
}

// =============================================================== WFT2
struct glsu_wft2 {
  WffPlan p;
  int type = GLSU_WFF;
  WfrBufs wfr;              // WFR mode only
  float* d_wfr_out = nullptr;  // WFR host form: staging for the eight maps
  hipStream_t s = nullptr;
  bool own = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float2* d_in = nullptr;   // staging for the host-pointer form
  float2* d_out = nullptr;
};

extern "C" int glsu_wft2_create_tuned(glsu_wft2** out, int W, int H, int type,
                                      const glsu_wff_params* pr, void* stream,
                                      const glsu_tuning* t) {
  if (!out || !pr) return fail(GLSU_ERR_ARG, "null argument");
  *out = nullptr;
  if (type != GLSU_WFF && type != GLSU_WFR) return fail(GLSU_ERR_ARG, "unknown WFT type %d", type);
  {
    int px = t ? t->planx : -1, py = t ? t->plany : -1;
    if (int rc = validate(W, H, pr, &px, &py)) return rc;
    if (int rc = validate_tuning(t)) return rc;
    if (type == GLSU_WFR)
      if (int rc = validate_wfr(W, H, pr)) return rc;
  }
  if (int rc = check_device()) return rc;
  auto* h = new glsu_wft2();
  h->type = type;
  int rc = plan_common(h->p, W, H, pr, t);
  if (rc) { delete h; return rc; }
  if (type == GLSU_WFR && (rc = wfr_alloc(h->p, h->wfr)) != GLSU_OK) {
    wfr_free(h->wfr);
    wff_plan_free(h->p);
    delete h;
    return rc;
  }
  rc = make_stream(stream, &h->s, &h->own);
  if (rc) { wff_plan_free(h->p); delete h; return rc; }
  hipEventCreate(&h->e0);
  hipEventCreate(&h->e1);
  *out = h;
  return GLSU_OK;
}

extern "C" int glsu_wft2_create(glsu_wft2** out, int W, int H, int type, const glsu_wff_params* p,
                                void* stream) {
  return glsu_wft2_create_tuned(out, W, H, type, p, stream, nullptr);
}

static int wft2_enqueue(glsu_wft2* h, const float2* f, float2* out) {
  if (!h->p.thr_ready) {
    HIP_TRY(launch_thr(h->p, f, h->s));
    h->p.thr_ready = true;
  }
  HIP_TRY(launch_rows_fwd(h->p, f, h->s));
  HIP_TRY(launch_cols(h->p, h->s));
  HIP_TRY(launch_rows_inv_filtered(h->p, out, h->s));
  return GLSU_OK;
}

extern "C" int glsu_wft2_compute(glsu_wft2* h, const void* d_f, void* d_out, double* ms) {
  if (!h || !d_f || !d_out) return fail(GLSU_ERR_ARG, "null argument");
  if (h->type != GLSU_WFF)
    return fail(GLSU_ERR_ARG, "this handle is in WFR mode: use glsu_wft2_compute_wfr");
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  if (int rc = wft2_enqueue(h, (const float2*)d_f, (float2*)d_out)) return rc;
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_wft2_compute_host(glsu_wft2* h, const float* hf, float* hout, double* ms) {
  if (!h || !hf || !hout) return fail(GLSU_ERR_ARG, "null argument");
  if (h->type != GLSU_WFF)
    return fail(GLSU_ERR_ARG, "this handle is in WFR mode: use glsu_wft2_compute_wfr_host");
  const size_t bytes = sizeof(float2) * (size_t)h->p.W * h->p.H;
  if (!h->d_in) {
    HIP_TRY(hipMalloc((void**)&h->d_in, bytes));
    HIP_TRY(hipMalloc((void**)&h->d_out, bytes));
  }
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  HIP_TRY(hipMemcpyAsync(h->d_in, hf, bytes, hipMemcpyHostToDevice, h->s));
  if (int rc = wft2_enqueue(h, h->d_in, h->d_out)) return rc;
  HIP_TRY(hipMemcpyAsync(hout, h->d_out, bytes, hipMemcpyDeviceToHost, h->s));
  if (ms) HIP_TRY(hipEventRecord(h->e1, h->s));
  HIP_TRY(hipStreamSynchronize(h->s));
  if (ms) {
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

// ---- WFR mode ---------------------------------------------------------------
static int wfr_enqueue(glsu_wft2* h, const float2* f, const WfrOut& z) {
  HIP_TRY(launch_rows_fwd(h->p, f, h->s));
  HIP_TRY(launch_wfr(h->p, h->wfr, z, h->s));
  return GLSU_OK;
}

static int wfr_out_from(const glsu_wfr_results* r, WfrOut* z) {
  if (!r || !r->wx || !r->wy || !r->phase || !r->phase_comp || !r->b || !r->r || !r->cxx ||
      !r->cyy)
    return fail(GLSU_ERR_ARG, "all eight WFR result pointers are required");
  *z = WfrOut{r->wx, r->wy, r->phase, r->phase_comp, r->b, r->r, r->cxx, r->cyy};
  return GLSU_OK;
}

extern "C" int glsu_wft2_compute_wfr(glsu_wft2* h, const void* d_f, const glsu_wfr_results* z,
                                     double* ms) {
  if (!h || !d_f) return fail(GLSU_ERR_ARG, "null argument");
  if (h->type != GLSU_WFR) return fail(GLSU_ERR_ARG, "this handle is in WFF mode");
  WfrOut o;
  if (int rc = wfr_out_from(z, &o)) return rc;
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  if (int rc = wfr_enqueue(h, (const float2*)d_f, o)) return rc;
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_wft2_compute_wfr_host(glsu_wft2* h, const float* hf, float* hout,
                                          double* ms) {
  if (!h || !hf || !hout) return fail(GLSU_ERR_ARG, "null argument");
  if (h->type != GLSU_WFR) return fail(GLSU_ERR_ARG, "this handle is in WFF mode");
  const size_t n = (size_t)h->p.W * h->p.H;
  if (!h->d_in) HIP_TRY(hipMalloc((void**)&h->d_in, sizeof(float2) * n));
  if (!h->d_wfr_out) HIP_TRY(hipMalloc((void**)&h->d_wfr_out, sizeof(float) * 8 * n));
  float* b = h->d_wfr_out;
  const WfrOut o{b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n, b + 7 * n};
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  HIP_TRY(hipMemcpyAsync(h->d_in, hf, sizeof(float2) * n, hipMemcpyHostToDevice, h->s));
  if (int rc = wfr_enqueue(h, h->d_in, o)) return rc;
  HIP_TRY(hipMemcpyAsync(hout, b, sizeof(float) * 8 * n, hipMemcpyDeviceToHost, h->s));
  if (ms) HIP_TRY(hipEventRecord(h->e1, h->s));
  HIP_TRY(hipStreamSynchronize(h->s));
  if (ms) {
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

// Average time of one stage over reps back-to-back launches on the handle's
// stream, after one full transform of d_f (which leaves the intermediate
// planes in place): 1 the row kernel, 2 the column kernel (WFF or WFR), 3 the
// rest (WFF: the inverse rows; WFR: ridge + curvature, writing into scratch
// maps).  No reference counterpart (measurement).
extern "C" int glsu_wft2_time_kernel(glsu_wft2* h, const void* d_f, int stage, int reps,
                                     double* ms) {
  if (!h || !d_f || !ms || reps < 1 || stage < 1 || stage > 3)
    return fail(GLSU_ERR_ARG, "invalid argument");
  const float2* f = (const float2*)d_f;
  const size_t n = (size_t)h->p.W * h->p.H;
  float2* out = nullptr;
  float* maps = nullptr;
  WfrOut z{};
  if (h->type == GLSU_WFF) {
    HIP_TRY(hipMalloc((void**)&out, sizeof(float2) * n));
    if (int rc = wft2_enqueue(h, f, out)) { hipFree(out); return rc; }
  } else {
    HIP_TRY(hipMalloc((void**)&maps, sizeof(float) * 8 * n));
    z = WfrOut{maps, maps + n, maps + 2 * n, maps + 3 * n, maps + 4 * n, maps + 5 * n,
               maps + 6 * n, maps + 7 * n};
    if (int rc = wfr_enqueue(h, f, z)) { hipFree(maps); return rc; }
  }
  hipError_t e = hipEventRecord(h->e0, h->s);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    if (stage == 1) e = launch_rows_fwd(h->p, f, h->s);
    else if (stage == 2) e = h->type == GLSU_WFF ? launch_cols(h->p, h->s) : launch_wfr_cols(h->p, h->wfr, h->s);
    else e = h->type == GLSU_WFF ? launch_rows_inv_filtered(h->p, out, h->s) : launch_wfr_post(h->p, h->wfr, z, h->s);
  }
  if (e == hipSuccess) e = hipEventRecord(h->e1, h->s);
  if (e == hipSuccess) e = hipEventSynchronize(h->e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, h->e0, h->e1);
  hipFree(out);
  hipFree(maps);
  if (e != hipSuccess) return fail(GLSU_ERR_DEVICE, "timing failed: %s", hipGetErrorString(e));
  *ms = t / reps;
  return GLSU_OK;
}

static double read_thr(const WffPlan& p, hipStream_t s) {
  if (!p.thr_ready) return std::numeric_limits<double>::quiet_NaN();
  float t = 0;
  hipMemcpyAsync(&t, p.d_thr, sizeof(float), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  return t;
}

static void get_tuning(const WffPlan& p, glsu_tuning* t) {
  t->planx = p.planx;
  t->plany = p.plany;
  t->cols_waves = p.cols_waves;
  t->rows_waves = p.rows_waves;
  t->ls_tile = p.ls_tile;
}

extern "C" int glsu_wft2_get_tuning(glsu_wft2* h, glsu_tuning* t) {
  if (!h || !t) return fail(GLSU_ERR_ARG, "null argument");
  get_tuning(h->p, t);
  return GLSU_OK;
}

extern "C" int glsu_wft2_info(glsu_wft2* h, int* nx, int* ny, int* sx, int* sy, int* nwx, int* nwy,
                              double* thr) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  if (nx) *nx = h->p.Nx;
  if (ny) *ny = h->p.Ny;
  if (sx) *sx = h->p.sx;
  if (sy) *sy = h->p.sy;
  if (nwx) *nwx = h->p.nwx;
  if (nwy) *nwy = h->p.nwy;
  if (thr) *thr = read_thr(h->p, h->s);
  return GLSU_OK;
}

extern "C" int glsu_wft2_destroy(glsu_wft2* h) {
  if (!h) return GLSU_OK;
  hipStreamSynchronize(h->s);
  wff_plan_free(h->p);
  wfr_free(h->wfr);
  hipFree(h->d_wfr_out);
  hipFree(h->d_in);
  hipFree(h->d_out);
  hipEventDestroy(h->e0);
  hipEventDestroy(h->e1);
  if (h->own) hipStreamDestroy(h->s);
  delete h;
  return GLSU_OK;
}

// =============================================================== DPRA
struct GraphKey {
  const void* frames;
  int nframes, dtype;
  float* dphi;
  const float* phiref;
  bool thr_ready;
  bool fresh;
  bool operator<(const GraphKey& o) const {
    return std::tie(frames, nframes, dtype, dphi, phiref, thr_ready, fresh) <
           std::tie(o.frames, o.nframes, o.dtype, o.dphi, o.phiref, o.thr_ready, o.fresh);
  }
};

struct GraphEntry {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  // host-side pointer state at the end of the captured sequence
  float *phiref, *phicurr, *dref_hi, *dref_lo, *dphi_hi, *dphi_lo;
  bool fresh;
};

struct glsu_dpra {
  int W = 0, H = 0, rr = 1;
  WffPlan p;
  hipStream_t s = nullptr;
  bool own = false;
  float *phiref = nullptr, *phicurr = nullptr;
  float *dref_hi = nullptr, *dref_lo = nullptr, *dphi_hi = nullptr, *dphi_lo = nullptr;
  float2* z = nullptr;
  bool fresh = true;  // phicurr/dphi hold values not yet taken into the reference
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  float* d_dphi_stage = nullptr;
  hipStream_t cap = nullptr;  // capture stream for glsu_dpra_run_graph
  std::map<GraphKey, GraphEntry> graphs;
  // learned path (glsu_dpra_frame_learned)
  float2* d_net = nullptr;                      // network output, H*W
  float2* d_fwin = nullptr;  // spot-check window, filtered
  size_t win_cap = 0;
  std::map<std::pair<int, int>, WffPlan> wins;  // WFF2 plans by window size
  int* d_flags = nullptr;                       // [fallback, !fallback] of the last frame
  float* d_stats = nullptr;                     // frames, fallbacks, sum rms, max rms, last rms
  float* h_res = nullptr;                       // pinned: last rms, fallback flag
};

static int dpra_alloc(glsu_dpra* h) {
  const size_t n = (size_t)h->W * h->H;
  HIP_TRY(hipMalloc((void**)&h->phiref, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->phicurr, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->dref_hi, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->dref_lo, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->dphi_hi, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->dphi_lo, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&h->z, n * sizeof(float2)));
  // DPRA_CPU ctor: phicurr = 0, dphiref = 0 (DPRA/dpra_cpu.cpp:17-20)
  HIP_TRY(hipMemset(h->phicurr, 0, n * sizeof(float)));
  HIP_TRY(hipMemset(h->dref_hi, 0, n * sizeof(float)));
  HIP_TRY(hipMemset(h->dref_lo, 0, n * sizeof(float)));
  HIP_TRY(hipMemset(h->dphi_hi, 0, n * sizeof(float)));
  HIP_TRY(hipMemset(h->dphi_lo, 0, n * sizeof(float)));
  return GLSU_OK;
}

static void dpra_free(glsu_dpra* h) {
  for (auto& kv : h->graphs) {
    hipGraphExecDestroy(kv.second.x);
    hipGraphDestroy(kv.second.g);
  }
  h->graphs.clear();
  hipFree(h->phiref); hipFree(h->phicurr);
  hipFree(h->dref_hi); hipFree(h->dref_lo); hipFree(h->dphi_hi); hipFree(h->dphi_lo);
  hipFree(h->z); hipFree(h->d_stage); hipFree(h->d_dphi_stage);
  hipFree(h->d_net); hipFree(h->d_fwin); hipFree(h->d_flags);
  hipFree(h->d_stats);
  if (h->h_res) hipHostFree(h->h_res);
  for (auto& kv : h->wins) wff_plan_free(kv.second);
  wff_plan_free(h->p);
  if (h->e0) hipEventDestroy(h->e0);
  if (h->e1) hipEventDestroy(h->e1);
  if (h->own && h->s) hipStreamDestroy(h->s);
  if (h->cap) hipStreamDestroy(h->cap);
}

// phase values are only used through cos/sin and differences; keep them in
// (-pi, pi] so FP32 storage keeps full relative precision
static float wrap_d(double v) { return (float)std::atan2(std::sin(v), std::cos(v)); }

template <typename PT>
static int dpra_create_impl(glsu_dpra** out, const PT* phi0, int W, int H, int rr,
                            const glsu_wff_params* pr, void* stream, const glsu_tuning* t) {
  if (!out || !phi0) return fail(GLSU_ERR_ARG, "null argument");
  *out = nullptr;
  if (rr < 1) return fail(GLSU_ERR_ARG, "reference update rate must be >= 1");
  glsu_wff_params def;
  if (!pr) {
    glsu_wff_params_dpra_reference(&def);
    pr = &def;
  }
  {
    int px = t ? t->planx : -1, py = t ? t->plany : -1;
    if (int rc = validate(W, H, pr, &px, &py)) return rc;
    if (int rc = validate_tuning(t)) return rc;
  }
  if (int rc = check_device()) return rc;
  auto* h = new glsu_dpra();
  h->W = W; h->H = H; h->rr = rr;
  int rc = plan_common(h->p, W, H, pr, t);
  if (rc) { delete h; return rc; }
  rc = make_stream(stream, &h->s, &h->own);
  if (!rc) rc = dpra_alloc(h);
  if (rc) { dpra_free(h); delete h; return rc; }
  std::vector<float> ph((size_t)W * H);
  for (size_t i = 0; i < ph.size(); ++i) ph[i] = wrap_d((double)phi0[i]);
  if (hipMemcpy(h->phiref, ph.data(), ph.size() * sizeof(float), hipMemcpyHostToDevice) !=
      hipSuccess) {
    dpra_free(h);
    delete h;
    return fail(GLSU_ERR_DEVICE, "upload of phi0 failed");
  }
  hipEventCreate(&h->e0);
  hipEventCreate(&h->e1);
  *out = h;
  return GLSU_OK;
}

extern "C" int glsu_dpra_create(glsu_dpra** out, const double* phi0, int W, int H, int rr,
                                const glsu_wff_params* p, void* stream) {
  return dpra_create_impl(out, phi0, W, H, rr, p, stream, nullptr);
}
extern "C" int glsu_dpra_create_f32(glsu_dpra** out, const float* phi0, int W, int H, int rr,
                                    const glsu_wff_params* p, void* stream) {
  return dpra_create_impl(out, phi0, W, H, rr, p, stream, nullptr);
}
extern "C" int glsu_dpra_create_tuned(glsu_dpra** out, const double* phi0, int W, int H, int rr,
                                      const glsu_wff_params* p, void* stream,
                                      const glsu_tuning* t) {
  return dpra_create_impl(out, phi0, W, H, rr, p, stream, t);
}

// one frame: LS -> (threshold on first call) -> rows fwd -> columns -> rows inv + update
static int dpra_enqueue(glsu_dpra* h, const void* frame, int dtype, float* d_dphi,
                        hipStream_t st) {
  WffPlan& p = h->p;
  HIP_TRY(launch_dpra_ls(frame, dtype, h->phiref, h->z, h->W, h->H, h->W, st, h->p.ls_tile));
  if (!p.thr_ready) {
    HIP_TRY(launch_thr(p, h->z, st));
    p.thr_ready = true;
  }
  HIP_TRY(launch_rows_fwd(p, h->z, st));
  HIP_TRY(launch_cols(p, st));
  UpdateArgs u{h->phiref, h->dref_hi, h->dref_lo, h->phicurr, h->dphi_hi, h->dphi_lo, d_dphi,
               nullptr};
  HIP_TRY(launch_rows_inv_update(p, u, st));
  h->fresh = true;
  return GLSU_OK;
}

// update_ref_phi, DPRA/dpra_cpu.cpp:254-258: phiref <- phicurr, dphiref <- dphi.
// Realised as a buffer swap; a repeated call without a new frame is a no-op,
// which is what repeated copies of unchanged buffers amount to.
static void dpra_swap(glsu_dpra* h) {
  if (!h->fresh) return;
  std::swap(h->phiref, h->phicurr);
  std::swap(h->dref_hi, h->dphi_hi);
  std::swap(h->dref_lo, h->dphi_lo);
  h->fresh = false;
}

extern "C" int glsu_dpra_frame(glsu_dpra* h, const void* d_frame, int dtype, float* d_dphi,
                               float* d_phicurr, double* ms) {
  if (!h || !d_frame) return fail(GLSU_ERR_ARG, "null argument");
  if (!elem_size(dtype)) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  if (int rc = dpra_enqueue(h, d_frame, dtype, d_dphi, h->s)) return rc;
  if (d_phicurr)
    HIP_TRY(hipMemcpyAsync(d_phicurr, h->phicurr, sizeof(float) * (size_t)h->W * h->H,
                           hipMemcpyDeviceToDevice, h->s));
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_frame_host(glsu_dpra* h, const void* hf, int dtype, float* hdphi,
                                    double* ms) {
  if (!h || !hf) return fail(GLSU_ERR_ARG, "null argument");
  const int es = elem_size(dtype);
  if (!es) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  const size_t n = (size_t)h->W * h->H;
  if (h->stage_bytes < n * es) {
    hipFree(h->d_stage);
    HIP_TRY(hipMalloc(&h->d_stage, n * es));
    h->stage_bytes = n * es;
  }
  if (!h->d_dphi_stage) HIP_TRY(hipMalloc((void**)&h->d_dphi_stage, n * sizeof(float)));
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  HIP_TRY(hipMemcpyAsync(h->d_stage, hf, n * es, hipMemcpyHostToDevice, h->s));
  if (int rc = dpra_enqueue(h, h->d_stage, dtype, h->d_dphi_stage, h->s)) return rc;
  if (hdphi)
    HIP_TRY(hipMemcpyAsync(hdphi, h->d_dphi_stage, n * sizeof(float), hipMemcpyDeviceToHost, h->s));
  if (ms) HIP_TRY(hipEventRecord(h->e1, h->s));
  HIP_TRY(hipStreamSynchronize(h->s));
  if (ms) {
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_update_ref(glsu_dpra* h) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  dpra_swap(h);
  return GLSU_OK;
}

static int dpra_run_enqueue(glsu_dpra* h, const void* frames, int nframes, int dtype,
                            float* dphi_all, hipStream_t st) {
  const size_t n = (size_t)h->W * h->H;
  const int es = elem_size(dtype);
  for (int i = 0; i < nframes; ++i) {
    const char* f = (const char*)frames + (size_t)i * n * es;
    if (int rc = dpra_enqueue(h, f, dtype, dphi_all ? dphi_all + (size_t)i * n : nullptr, st))
      return rc;
    if (i % h->rr == 0) dpra_swap(h);
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_run(glsu_dpra* h, const void* frames, int nframes, int dtype,
                             float* dphi_all, double* ms) {
  if (!h || nframes < 0 || (!frames && nframes > 0)) return fail(GLSU_ERR_ARG, "invalid argument");
  if (!elem_size(dtype)) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  if (int rc = dpra_run_enqueue(h, frames, nframes, dtype, dphi_all, h->s)) return rc;
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_run_graph(glsu_dpra* h, const void* frames, int nframes, int dtype,
                                   float* dphi_all, double* ms) {
  if (!h || nframes < 0 || (!frames && nframes > 0)) return fail(GLSU_ERR_ARG, "invalid argument");
  if (!elem_size(dtype)) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  GraphKey key{frames, nframes, dtype, dphi_all, h->phiref, h->p.thr_ready, h->fresh};
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    GraphEntry ge;
    // capture on a private stream (the null stream cannot be captured); the
    // graph is launched on the handle's stream below
    if (!h->cap) HIP_TRY(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
    int rc = dpra_run_enqueue(h, frames, nframes, dtype, dphi_all, h->cap);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(h->cap, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (ee != hipSuccess) return fail(GLSU_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(ee));
    ge.g = g;
    HIP_TRY(hipGraphInstantiate(&ge.x, g, nullptr, nullptr, 0));
    ge.phiref = h->phiref; ge.phicurr = h->phicurr;
    ge.dref_hi = h->dref_hi; ge.dref_lo = h->dref_lo;
    ge.dphi_hi = h->dphi_hi; ge.dphi_lo = h->dphi_lo;
    ge.fresh = h->fresh;
    it = h->graphs.emplace(key, ge).first;
  } else {
    // replay ends in the same host pointer state the capture ended in
    h->phiref = it->second.phiref; h->phicurr = it->second.phicurr;
    h->dref_hi = it->second.dref_hi; h->dref_lo = it->second.dref_lo;
    h->dphi_hi = it->second.dphi_hi; h->dphi_lo = it->second.dphi_lo;
    h->fresh = it->second.fresh;
    h->p.thr_ready = true;
  }
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  HIP_TRY(hipGraphLaunch(it->second.x, h->s));
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

// Per-kernel device times over a frame loop (events between the kernels of
// each frame).  ms[0..3] receive the summed times of the least-squares, row
// forward, column and row inverse + update kernels.  Diagnostics only.
static int dpra_profile_impl(glsu_dpra* h, const void* frames, int nframes, int dtype,
                             float* dphi_all, double* ms);

extern "C" int glsu_dpra_profile(glsu_dpra* h, const void* frames, int nframes, int dtype,
                                 float* dphi_all, double* ms) {
  return dpra_profile_impl(h, frames, nframes, dtype, dphi_all, ms);
}

// The same, counting the column transforms skipped because a column's
// thresholded signal was zero on every image row (a device counter the
// column kernel increments only in this call).
extern "C" int glsu_dpra_profile_skips(glsu_dpra* h, const void* frames, int nframes, int dtype,
                                       float* dphi_all, double* ms, unsigned long long* skipped,
                                       unsigned long long* transforms) {
  if (!h || !skipped || !transforms) return fail(GLSU_ERR_ARG, "invalid argument");
  unsigned int* cnt = nullptr;
  HIP_TRY(hipMalloc((void**)&cnt, sizeof(unsigned int)));
  int rc = GLSU_OK;
  if (hipMemsetAsync(cnt, 0, sizeof(unsigned int), h->s) != hipSuccess) {
    rc = fail(GLSU_ERR_DEVICE, "counter reset failed");
  } else {
    h->p.d_nskip = cnt;
    rc = dpra_profile_impl(h, frames, nframes, dtype, dphi_all, ms);
    h->p.d_nskip = nullptr;
  }
  unsigned int v = 0;
  if (!rc && hipMemcpy(&v, cnt, sizeof v, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(GLSU_ERR_DEVICE, "counter read failed");
  hipFree(cnt);
  if (rc) return rc;
  *skipped = v;
  *transforms = (unsigned long long)nframes * h->p.nwx * h->W * h->p.nwy;
  return GLSU_OK;
}

// Events are recorded for up to kProfBatch frames before the host waits for
// them, so the kernels run back to back as in the chain (a host wait after
// every frame would leave the device idle between frames and slow the next
// kernels' start).
static int dpra_profile_impl(glsu_dpra* h, const void* frames, int nframes, int dtype,
                             float* dphi_all, double* ms) {
  if (!h || !frames || !ms || nframes < 0) return fail(GLSU_ERR_ARG, "invalid argument");
  const int es = elem_size(dtype);
  if (!es) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  const size_t n = (size_t)h->W * h->H;
  constexpr int kProfBatch = 64;
  std::vector<hipEvent_t> ev(5 * kProfBatch, nullptr);
  int rc = GLSU_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = fail(GLSU_ERR_DEVICE, "hipEventCreate failed");
  for (int k = 0; k < 4; ++k) ms[k] = 0;
  WffPlan& p = h->p;
  auto run = [&]() -> int {
    for (int i0 = 0; i0 < nframes; i0 += kProfBatch) {
      const int nb = std::min(kProfBatch, nframes - i0);
      for (int b = 0; b < nb; ++b) {
        const int i = i0 + b;
        hipEvent_t* e = &ev[5 * b];
        const char* f = (const char*)frames + (size_t)i * n * es;
        HIP_TRY(hipEventRecord(e[0], h->s));
        HIP_TRY(launch_dpra_ls(f, dtype, h->phiref, h->z, h->W, h->H, h->W, h->s, h->p.ls_tile));
        if (!p.thr_ready) {
          HIP_TRY(launch_thr(p, h->z, h->s));
          p.thr_ready = true;
        }
        HIP_TRY(hipEventRecord(e[1], h->s));
        HIP_TRY(launch_rows_fwd(p, h->z, h->s));
        HIP_TRY(hipEventRecord(e[2], h->s));
        HIP_TRY(launch_cols(p, h->s));
        HIP_TRY(hipEventRecord(e[3], h->s));
        UpdateArgs u{h->phiref, h->dref_hi, h->dref_lo, h->phicurr, h->dphi_hi, h->dphi_lo,
                     dphi_all ? dphi_all + (size_t)i * n : nullptr, nullptr};
        HIP_TRY(launch_rows_inv_update(p, u, h->s));
        HIP_TRY(hipEventRecord(e[4], h->s));
        h->fresh = true;
        if (i % h->rr == 0) dpra_swap(h);
      }
      HIP_TRY(hipEventSynchronize(ev[5 * (nb - 1) + 4]));
      for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 4; ++k) {
          float t = 0;
          HIP_TRY(hipEventElapsedTime(&t, ev[5 * b + k], ev[5 * b + k + 1]));
          ms[k] += t;
        }
    }
    return GLSU_OK;
  };
  if (!rc) rc = run();
  for (auto& e : ev)
    if (e) hipEventDestroy(e);
  return rc;
}

extern "C" int glsu_dpra_time_kernel(glsu_dpra* h, int stage, int reps, double* ms) {
  if (!h || !ms || reps < 1 || stage < 1 || stage > 3)
    return fail(GLSU_ERR_ARG, "invalid argument (stage 1..3, reps >= 1)");
  WffPlan& p = h->p;
  if (!p.thr_ready) return fail(GLSU_ERR_ARG, "run one frame before timing a kernel");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // every stage below is idempotent on the handle's buffers (no state swap)
  UpdateArgs u{h->phiref, h->dref_hi, h->dref_lo, h->phicurr, h->dphi_hi, h->dphi_lo, nullptr,
               nullptr};
  auto one = [&]() -> hipError_t {
    if (stage == 1) return launch_rows_fwd(p, h->z, h->s);
    if (stage == 2) return launch_cols(p, h->s);
    return launch_rows_inv_update(p, u, h->s);
  };
  HIP_TRY(one());  // warm
  HIP_TRY(hipEventRecord(e0, h->s));
  for (int i = 0; i < reps; ++i) HIP_TRY(one());
  HIP_TRY(hipEventRecord(e1, h->s));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return GLSU_OK;
}

extern "C" int glsu_dpra_get_state(glsu_dpra* h, float* phiref, float* dphiref) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  const size_t n = (size_t)h->W * h->H;
  HIP_TRY(hipStreamSynchronize(h->s));
  if (phiref) HIP_TRY(hipMemcpy(phiref, h->phiref, n * sizeof(float), hipMemcpyDeviceToHost));
  if (dphiref) {
    std::vector<float> hi(n), lo(n);
    HIP_TRY(hipMemcpy(hi.data(), h->dref_hi, n * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lo.data(), h->dref_lo, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) dphiref[i] = hi[i] + lo[i];
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_set_state(glsu_dpra* h, const float* phiref, const float* dphiref) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  const size_t n = (size_t)h->W * h->H;
  HIP_TRY(hipStreamSynchronize(h->s));
  if (phiref) {
    std::vector<float> w(n);
    for (size_t i = 0; i < n; ++i) w[i] = wrap_d((double)phiref[i]);
    HIP_TRY(hipMemcpy(h->phiref, w.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  if (dphiref) {
    HIP_TRY(hipMemcpy(h->dref_hi, dphiref, n * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(h->dref_lo, 0, n * sizeof(float)));
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_copy_z(glsu_dpra* h, void* d_z) {
  if (!h || !d_z) return fail(GLSU_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(d_z, h->z, sizeof(float2) * (size_t)h->W * h->H,
                         hipMemcpyDeviceToDevice, h->s));
  return GLSU_OK;
}

extern "C" int glsu_dpra_thr(glsu_dpra* h, double* thr) {
  if (!h || !thr) return fail(GLSU_ERR_ARG, "null argument");
  *thr = read_thr(h->p, h->s);
  return GLSU_OK;
}

extern "C" int glsu_dpra_get_tuning(glsu_dpra* h, glsu_tuning* t) {
  if (!h || !t) return fail(GLSU_ERR_ARG, "null argument");
  get_tuning(h->p, t);
  return GLSU_OK;
}

extern "C" int glsu_dpra_destroy(glsu_dpra* h) {
  if (!h) return GLSU_OK;
  hipStreamSynchronize(h->s);
  dpra_free(h);
  delete h;
  return GLSU_OK;
}

extern "C" int glsu_gen_frames(float* d_out, int W, int H, int t0, int T, uint64_t seed, double amp0,
                               double amp1, double carrier, double noise, void* stream) {
  if (!d_out || W < 1 || H < 1 || T < 0) return fail(GLSU_ERR_ARG, "invalid argument");
  if (int rc = check_device()) return rc;
  HIP_TRY(launch_gen_frames(d_out, W, H, t0, T, seed, amp0, amp1, carrier, noise,
                            (hipStream_t)stream));
  return GLSU_OK;
}

// ============================================================ DPRA batch
// Many independent sequences of uneven sizes in one set of persistent
// launches per frame: one least-squares launch over every image, and per FFT
// plan class (images sharing planx/plany/launch shapes) one launch per WFF2
// stage, all driven by a device table of images and a device frame counter,
// so one captured graph serves every frame.  Each image is computed by the
// same per-item code as a single-sequence handle with the same tuning, so the
// results are bit-identical to glsu_dpra_run on that handle.
namespace {
struct BatchImage {
  int W = 0, H = 0, rr = 1;
  float *phi[2] = {nullptr, nullptr}, *dhi[2] = {nullptr, nullptr}, *dlo[2] = {nullptr, nullptr};
  float2 *z = nullptr, *A = nullptr, *D = nullptr;
  size_t plane = 0;
  float *thr = nullptr, *red = nullptr;
  int seq = 0;  // index in the caller's list
};
struct BatchCls {
  WffPlan p;  // tables only (spectra, twiddles), plan ids and launch shapes
  int b = 0, e = 0;
  BatchClass bc{};
};
}  // namespace

struct glsu_dpra_batch {
  int n = 0;
  hipStream_t s = nullptr, cap = nullptr;
  bool own = false;
  std::vector<BatchImage> img;  // table order
  std::vector<int> pos;         // caller index -> table index
  std::vector<BatchCls> cls;
  BatchSeg* d_tab = nullptr;
  BatchSeg* h_tab = nullptr;    // pinned staging of the table
  int* d_frame = nullptr;
  int* h_frame = nullptr;       // pinned
  int ntiles = 0, grid_ls = 0;
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;  // by frame dtype
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

static void batch_free(glsu_dpra_batch* h) {
  if (h->s) hipStreamSynchronize(h->s);
  for (auto& kv : h->graphs) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  for (auto& im : h->img) {
    for (int k = 0; k < 2; ++k) {
      hipFree(im.phi[k]); hipFree(im.dhi[k]); hipFree(im.dlo[k]);
    }
    hipFree(im.z); hipFree(im.A); hipFree(im.D); hipFree(im.thr); hipFree(im.red);
  }
  for (auto& c : h->cls) wff_plan_free(c.p);
  hipFree(h->d_tab); hipFree(h->d_frame);
  if (h->h_tab) hipHostFree(h->h_tab);
  if (h->h_frame) hipHostFree(h->h_frame);
  if (h->e0) hipEventDestroy(h->e0);
  if (h->e1) hipEventDestroy(h->e1);
  if (h->own && h->s) hipStreamDestroy(h->s);
  if (h->cap) hipStreamDestroy(h->cap);
}

static int batch_alloc_image(BatchImage& im, const WffPlan& cp, const double* phi0, double thr) {
  const size_t n = (size_t)im.W * im.H;
  const int wb = (im.W + kTileCols - 1) / kTileCols;
  im.plane = (size_t)wb * kTileCols * im.H;
  for (int k = 0; k < 2; ++k) {
    HIP_TRY(hipMalloc((void**)&im.phi[k], n * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&im.dhi[k], n * sizeof(float)));
    HIP_TRY(hipMalloc((void**)&im.dlo[k], n * sizeof(float)));
    HIP_TRY(hipMemset(im.dhi[k], 0, n * sizeof(float)));
    HIP_TRY(hipMemset(im.dlo[k], 0, n * sizeof(float)));
  }
  HIP_TRY(hipMemset(im.phi[1], 0, n * sizeof(float)));
  HIP_TRY(hipMalloc((void**)&im.z, n * sizeof(float2)));
  HIP_TRY(hipMalloc((void**)&im.A, im.plane * cp.nwx * sizeof(float2)));
  HIP_TRY(hipMalloc((void**)&im.D, im.plane * cp.nwx * sizeof(float2)));
  HIP_TRY(hipMalloc((void**)&im.thr, sizeof(float)));
  HIP_TRY(hipMalloc((void**)&im.red, 256 * sizeof(float)));
  const float t0 = thr >= 0 ? (float)thr : -1.f;
  HIP_TRY(hipMemcpy(im.thr, &t0, sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> ph(n);
  for (size_t i = 0; i < n; ++i) ph[i] = wrap_d(phi0[i]);
  HIP_TRY(hipMemcpy(im.phi[0], ph.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return GLSU_OK;
}

extern "C" int glsu_dpra_batch_create(glsu_dpra_batch** out, const glsu_seq_desc* seqs, int n,
                                      const glsu_wff_params* pr, void* stream,
                                      const glsu_tuning* tunings) {
  if (!out || !seqs || n < 1) return fail(GLSU_ERR_ARG, "invalid argument");
  *out = nullptr;
  glsu_wff_params def;
  if (!pr) {
    glsu_wff_params_dpra_reference(&def);
    pr = &def;
  }
  // host-side validation of every sequence and its plan choice
  // lx, ly: whether the image fits the live registers of its row / column
  // plan (the kernels compiled for them round slightly differently, and a
  // class shares one choice)
  struct Req { int px, py, cw, rw, lx, ly; };
  auto live_fit = [](int plan, int n) {
    for (const FftPlanInfo& q : fft_plans())
      if (q.id == plan) return n <= live_regs(q.E) * q.T ? 1 : 0;
    return 0;
  };
  std::vector<Req> req(n);
  for (int i = 0; i < n; ++i) {
    const glsu_seq_desc& d = seqs[i];
    if (!d.phi0) return fail(GLSU_ERR_ARG, "sequence %d: null phi0", i);
    if (d.rr < 1) return fail(GLSU_ERR_ARG, "sequence %d: reference update rate must be >= 1", i);
    const glsu_tuning* t = tunings ? &tunings[i] : nullptr;
    int px = t ? t->planx : -1, py = t ? t->plany : -1;
    if (int rc = validate(d.width, d.height, pr, &px, &py)) return rc;
    if (int rc = validate_tuning(t)) return rc;
    req[i] = {px, py, t ? t->cols_waves : -1, t ? t->rows_waves : -1, live_fit(px, d.width),
              live_fit(py, d.height)};
  }
  if (int rc = check_device()) return rc;
  auto* h = new glsu_dpra_batch();
  h->n = n;
  auto bail = [&](int rc) {
    batch_free(h);
    delete h;
    return rc;
  };
  if (int rc = make_stream(stream, &h->s, &h->own)) return bail(rc);
  // plan classes, in order of first appearance
  std::vector<int> cls_of(n, -1);
  std::vector<Req> keys;
  for (int i = 0; i < n; ++i) {
    int c = -1;
    for (size_t k = 0; k < keys.size(); ++k)
      if (keys[k].px == req[i].px && keys[k].py == req[i].py && keys[k].cw == req[i].cw &&
          keys[k].rw == req[i].rw && keys[k].lx == req[i].lx && keys[k].ly == req[i].ly)
        c = (int)k;
    if (c < 0) {
      c = (int)keys.size();
      keys.push_back(req[i]);
    }
    cls_of[i] = c;
  }
  h->cls.resize(keys.size());
  h->pos.assign(n, -1);
  for (size_t c = 0; c < keys.size(); ++c) {
    BatchCls& bc = h->cls[c];
    int first = -1;
    for (int i = 0; i < n; ++i)
      if (cls_of[i] == (int)c && first < 0) first = i;
    std::string err;
    int rc = wff_plan_init(bc.p, seqs[first].width, seqs[first].height, pr->sigma_x, pr->wxl,
                           pr->wxh, pr->wxi, pr->sigma_y, pr->wyl, pr->wyh, pr->wyi, pr->thr,
                           keys[c].px, keys[c].py, keys[c].cw, keys[c].rw, err, false);
    if (rc) return bail(fail(rc == -2 ? GLSU_ERR_ARG : GLSU_ERR_DEVICE, "%s", err.c_str()));
    bc.b = (int)h->img.size();
    for (int i = 0; i < n; ++i) {
      if (cls_of[i] != (int)c) continue;
      BatchImage im;
      im.W = seqs[i].width; im.H = seqs[i].height; im.rr = seqs[i].rr; im.seq = i;
      h->pos[i] = (int)h->img.size();
      h->img.push_back(im);
      if (int e = batch_alloc_image(h->img.back(), bc.p, seqs[i].phi0, pr->thr)) return bail(e);
    }
    bc.e = (int)h->img.size();
  }
  // table offsets and launch geometry
  const int nt = (int)h->img.size();
  if (hipHostMalloc((void**)&h->h_tab, sizeof(BatchSeg) * nt) != hipSuccess ||
      hipHostMalloc((void**)&h->h_frame, sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&h->d_tab, sizeof(BatchSeg) * nt) != hipSuccess ||
      hipMalloc((void**)&h->d_frame, sizeof(int)) != hipSuccess)
    return bail(fail(GLSU_ERR_DEVICE, "batch table allocation failed"));
  int tile = 0;
  for (auto& c : h->cls) {
    const WffPlan& p = c.p;
    const int ipb = cols_items_per_block(p.plany, p.cols_waves);
    int row = 0, blk = 0, maxW = 0, maxH = 0;
    for (int k = c.b; k < c.e; ++k) {
      BatchSeg& sg = h->h_tab[k];
      const BatchImage& im = h->img[k];
      maxW = std::max(maxW, im.W);
      maxH = std::max(maxH, im.H);
      sg = BatchSeg{};
      sg.W = im.W; sg.H = im.H; sg.rr = im.rr;
      sg.row0 = row; sg.col0 = blk; sg.tile0 = tile;
      sg.inv_n = (float)(1.0 / ((double)im.W * im.H));
      row += im.H;
      blk += p.nwx * ((im.W + ipb - 1) / ipb);
      tile += dpra_ls_tiles(im.W, im.H);
    }
    BatchClass& b = c.bc;
    b.tab = h->d_tab + c.b;
    b.nseg = c.e - c.b;
    b.nrows = row;
    b.ncolblk = blk;
    b.planx = p.planx; b.plany = p.plany;
    b.cols_waves = p.cols_waves; b.rows_waves = p.rows_waves;
    b.nwx = p.nwx; b.nwy = p.nwy; b.scale = p.scale;
    b.maxW = maxW; b.maxH = maxH;
    b.Hx = p.d_Hx; b.Hy = p.d_Hy; b.twx = p.d_twx; b.twy = p.d_twy;
    b.grid_rows = batch_grid_rows(p.planx, p.rows_waves, row);
    b.grid_cols = batch_grid_cols(p.plany, p.cols_waves, blk);
    if (b.grid_rows < 1 || b.grid_cols < 1)
      return bail(fail(GLSU_ERR_DEVICE, "occupancy query failed"));
  }
  h->ntiles = tile;
  h->grid_ls = batch_grid_ls(tile);
  if (h->grid_ls < 1) return bail(fail(GLSU_ERR_DEVICE, "occupancy query failed"));
  hipEventCreate(&h->e0);
  hipEventCreate(&h->e1);
  *out = h;
  return GLSU_OK;
}

// one frame of every sequence; the device frame counter selects frames,
// outputs and the reference/current buffers
static int batch_step(glsu_dpra_batch* h, int dtype, hipStream_t st) {
  HIP_TRY(launch_dpra_ls_batch(h->d_tab, (int)h->img.size(), h->ntiles, h->grid_ls, dtype,
                               h->d_frame, st));
  HIP_TRY(launch_thr_batch(h->d_tab, (int)h->img.size(), st));
  for (auto& c : h->cls) HIP_TRY(launch_rows_fwd_batch(c.bc, st));
  for (auto& c : h->cls) HIP_TRY(launch_cols_batch(c.bc, st));
  for (auto& c : h->cls) HIP_TRY(launch_rows_inv_batch(c.bc, h->d_frame, st));
  HIP_TRY(launch_frame_advance(h->d_frame, st));
  return GLSU_OK;
}

static int swaps_before(int t, int rr) { return t == 0 ? 0 : (t - 1) / rr + 1; }

extern "C" int glsu_dpra_batch_run(glsu_dpra_batch* h, const void* const* frames, int nframes,
                                   int dtype, float* const* dphi_all, int use_graph,
                                   double* ms) {
  if (!h || nframes < 0 || (!frames && nframes > 0)) return fail(GLSU_ERR_ARG, "invalid argument");
  if (!elem_size(dtype)) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  for (int i = 0; i < h->n; ++i)
    if (!frames[i]) return fail(GLSU_ERR_ARG, "sequence %d: null frames", i);
  if (nframes == 0) return GLSU_OK;
  HIP_TRY(hipStreamSynchronize(h->s));  // staging buffers are free
  for (size_t k = 0; k < h->img.size(); ++k) {
    const BatchImage& im = h->img[k];
    BatchSeg& sg = h->h_tab[k];
    sg.frames = frames[im.seq];
    sg.dphi_out = dphi_all ? dphi_all[im.seq] : nullptr;
    for (int b = 0; b < 2; ++b) {
      sg.phi[b] = im.phi[b]; sg.dhi[b] = im.dhi[b]; sg.dlo[b] = im.dlo[b];
    }
    sg.z = im.z; sg.A = im.A; sg.D = im.D; sg.plane = im.plane;
    sg.thr = im.thr; sg.red = im.red;
  }
  *h->h_frame = 0;
  HIP_TRY(hipMemcpyAsync(h->d_tab, h->h_tab, sizeof(BatchSeg) * h->img.size(),
                         hipMemcpyHostToDevice, h->s));
  HIP_TRY(hipMemcpyAsync(h->d_frame, h->h_frame, sizeof(int), hipMemcpyHostToDevice, h->s));
  if (ms) HIP_TRY(hipEventRecord(h->e0, h->s));
  if (use_graph) {
    auto it = h->graphs.find(dtype);
    if (it == h->graphs.end()) {
      if (!h->cap) HIP_TRY(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
      HIP_TRY(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
      int rc = batch_step(h, dtype, h->cap);
      hipGraph_t g = nullptr;
      hipError_t ee = hipStreamEndCapture(h->cap, &g);
      if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
      }
      if (ee != hipSuccess)
        return fail(GLSU_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(ee));
      hipGraphExec_t x = nullptr;
      HIP_TRY(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
      it = h->graphs.emplace(dtype, std::make_pair(g, x)).first;
    }
    for (int t = 0; t < nframes; ++t) HIP_TRY(hipGraphLaunch(it->second.second, h->s));
  } else {
    for (int t = 0; t < nframes; ++t)
      if (int rc = batch_step(h, dtype, h->s)) return rc;
  }
  if (ms) {
    HIP_TRY(hipEventRecord(h->e1, h->s));
    HIP_TRY(hipEventSynchronize(h->e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  // the next run starts from index 0 = reference again
  for (auto& im : h->img)
    if (swaps_before(nframes, im.rr) & 1) {
      std::swap(im.phi[0], im.phi[1]);
      std::swap(im.dhi[0], im.dhi[1]);
      std::swap(im.dlo[0], im.dlo[1]);
    }
  return GLSU_OK;
}

extern "C" int glsu_dpra_batch_info(glsu_dpra_batch* h, int* nclasses, int* ntiles) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  if (nclasses) *nclasses = (int)h->cls.size();
  if (ntiles) *ntiles = h->ntiles;
  return GLSU_OK;
}

extern "C" int glsu_dpra_batch_get_tuning(glsu_dpra_batch* h, int seq, glsu_tuning* t) {
  if (!h || !t || seq < 0 || seq >= h->n) return fail(GLSU_ERR_ARG, "invalid argument");
  const int k = h->pos[seq];
  for (auto& c : h->cls)
    if (k >= c.b && k < c.e) get_tuning(c.p, t);
  return GLSU_OK;
}

extern "C" int glsu_dpra_batch_get_state(glsu_dpra_batch* h, int seq, float* phiref,
                                         float* dphiref) {
  if (!h || seq < 0 || seq >= h->n) return fail(GLSU_ERR_ARG, "invalid argument");
  const BatchImage& im = h->img[h->pos[seq]];
  const size_t n = (size_t)im.W * im.H;
  HIP_TRY(hipStreamSynchronize(h->s));
  if (phiref) HIP_TRY(hipMemcpy(phiref, im.phi[0], n * sizeof(float), hipMemcpyDeviceToHost));
  if (dphiref) {
    std::vector<float> hi(n), lo(n);
    HIP_TRY(hipMemcpy(hi.data(), im.dhi[0], n * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lo.data(), im.dlo[0], n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) dphiref[i] = hi[i] + lo[i];
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_batch_destroy(glsu_dpra_batch* h) {
  if (!h) return GLSU_OK;
  batch_free(h);
  delete h;
  return GLSU_OK;
}

// ============================================================ learned frame
// One DPRA frame with the U-Net standing in for WFF2 (LS -> z -> net -> phase
// update), with a WFF2 spot check: the 64 x 64 tile at (tile_x, tile_y) is
// filtered exactly by WFF2 on its window (tile +- 2s, clipped to the image;
// each output pixel of WFF2 depends on its input within +-2s, so the window
// gives the full-image values on the tile), and when the RMS of the wrapped
// difference between the network's and WFF2's phase on the tile exceeds tol
// the frame is filtered by WFF2 instead.  tol < 0 skips the check.
extern "C" int glsu_dpra_frame_learned(glsu_dpra* h, glsu_unet* net, const void* frame,
                                       int dtype, float* d_dphi, double tol, int tile_x,
                                       int tile_y, double* residual, int* fallback) {
  if (!h || !net || !frame) return fail(GLSU_ERR_ARG, "null argument");
  if (!elem_size(dtype)) return fail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  const int W = h->W, H = h->H;
  if (tol >= 0 && (W < 64 || H < 64)) return fail(GLSU_ERR_ARG, "spot check needs >= 64x64");
  WffPlan& p = h->p;
  const size_t n = (size_t)W * H;
  if (!h->d_net) HIP_TRY(hipMalloc((void**)&h->d_net, n * sizeof(float2)));
  if (!h->d_flags) {
    HIP_TRY(hipMalloc((void**)&h->d_flags, 2 * sizeof(int)));
    HIP_TRY(hipMalloc((void**)&h->d_stats, 5 * sizeof(float)));
    HIP_TRY(hipMemset(h->d_stats, 0, 5 * sizeof(float)));
    HIP_TRY(hipHostMalloc((void**)&h->h_res, 2 * sizeof(float)));
  }
  HIP_TRY(launch_dpra_ls(frame, dtype, h->phiref, h->z, W, H, W, h->s, h->p.ls_tile));
  if (!p.thr_ready) {
    HIP_TRY(launch_thr(p, h->z, h->s));
    p.thr_ready = true;
  }
  if (int rc = unet_forward_on(net, h->z, h->d_net, H, W, h->s)) return rc;
  UpdateArgs u{h->phiref, h->dref_hi, h->dref_lo, h->phicurr, h->dphi_hi, h->dphi_lo, d_dphi,
               nullptr};
  if (tol < 0) {
    HIP_TRY(launch_net_update(h->d_net, u, n, h->s, nullptr));
    h->fresh = true;
    if (residual) *residual = -1.0;
    if (fallback) *fallback = 0;
    return GLSU_OK;
  }
  const int tx = std::min(std::max(tile_x, 0), W - 64), ty = std::min(std::max(tile_y, 0), H - 64);
  const int x0 = std::max(tx - 2 * p.sx, 0), x1 = std::min(tx + 64 + 2 * p.sx, W);
  const int y0 = std::max(ty - 2 * p.sy, 0), y1 = std::min(ty + 64 + 2 * p.sy, H);
  const int Ww = x1 - x0, Hw = y1 - y0;
  auto key = std::make_pair(Ww, Hw);
  auto it = h->wins.find(key);
  if (it == h->wins.end()) {
    WffPlan wp;
    std::string err;
    // same plan rule as the public constructors (row plans use 64 lanes)
    const int px = default_plan(Ww + p.sx, true), py = default_plan(Hw + p.sy, false);
    if (px < 0 || py < 0) return fail(GLSU_ERR_ARG, "no FFT plan for the spot-check window");
    int rc = wff_plan_init(wp, Ww, Hw, p.sigx, p.wxl, p.wxh, p.wxi, p.sigy, p.wyl, p.wyh, p.wyi,
                           -1.0, px, py, -1, -1, err);
    if (rc) return fail(GLSU_ERR_DEVICE, "spot-check plan: %s", err.c_str());
    it = h->wins.emplace(key, wp).first;
  }
  WffPlan& wp = it->second;
  if ((size_t)Ww * Hw > h->win_cap) {
    HIP_TRY(hipStreamSynchronize(h->s));
    hipFree(h->d_fwin);
    h->d_fwin = nullptr;
    h->win_cap = (size_t)Ww * Hw;
    HIP_TRY(hipMalloc((void**)&h->d_fwin, h->win_cap * sizeof(float2)));
  }
  // the window uses the full image's threshold, fixed once computed (it is
  // never recomputed on this handle): copied once per window plan
  if (!wp.thr_ready) {
    HIP_TRY(hipMemcpyAsync(wp.d_thr, p.d_thr, sizeof(float), hipMemcpyDeviceToDevice, h->s));
    wp.thr_ready = true;
  }
  // the window's rows are read in place from the full field (row pitch W)
  HIP_TRY(launch_rows_fwd(wp, h->z + (size_t)y0 * W + x0, h->s, nullptr, W));
  HIP_TRY(launch_cols(wp, h->s));
  HIP_TRY(launch_rows_inv_filtered(wp, h->d_fwin, h->s));
  HIP_TRY(launch_tile_residual(h->d_net, W, tx, ty, h->d_fwin, Ww, tx - x0, ty - y0, (float)tol,
                               h->d_flags, h->d_flags + 1, h->d_stats, h->s));
  // both paths are enqueued; the device flag lets exactly one of them run, so
  // the host never waits for the decision
  int* not_fb = h->d_flags + 1;
  HIP_TRY(launch_rows_fwd(p, h->z, h->s, not_fb));
  HIP_TRY(launch_cols(p, h->s, not_fb));
  HIP_TRY(launch_rows_inv_update(p, u, h->s, not_fb));
  HIP_TRY(launch_net_update(h->d_net, u, n, h->s, h->d_flags));
  h->fresh = true;
  if (residual || fallback) {
    HIP_TRY(hipMemcpyAsync(h->h_res, h->d_stats + 4, sizeof(float), hipMemcpyDeviceToHost, h->s));
    HIP_TRY(hipMemcpyAsync(h->h_res + 1, h->d_flags, sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIP_TRY(hipStreamSynchronize(h->s));
    int fb = 0;
    std::memcpy(&fb, h->h_res + 1, sizeof(int));
    if (residual) *residual = h->h_res[0];
    if (fallback) *fallback = fb;
  }
  return GLSU_OK;
}

extern "C" int glsu_dpra_learned_stats(glsu_dpra* h, int* frames, int* fallbacks,
                                       double* mean_rms, double* max_rms, int reset) {
  if (!h) return fail(GLSU_ERR_ARG, "null handle");
  float st[5] = {0, 0, 0, 0, 0};
  if (h->d_stats) {
    HIP_TRY(hipStreamSynchronize(h->s));
    HIP_TRY(hipMemcpy(st, h->d_stats, sizeof st, hipMemcpyDeviceToHost));
    if (reset) HIP_TRY(hipMemset(h->d_stats, 0, sizeof st));
  }
  if (frames) *frames = (int)st[0];
  if (fallbacks) *fallbacks = (int)st[1];
  if (mean_rms) *mean_rms = st[0] > 0 ? st[2] / st[0] : 0.0;
  if (max_rms) *max_rms = st[3];
  return GLSU_OK;
}

extern "C" int glsu_dpra_copy_net(glsu_dpra* h, void* d_out) {
  if (!h || !d_out) return fail(GLSU_ERR_ARG, "null argument");
  if (!h->d_net) return fail(GLSU_ERR_ARG, "no learned frame has run");
  HIP_TRY(hipMemcpyAsync(d_out, h->d_net, (size_t)h->W * h->H * sizeof(float2),
                         hipMemcpyDeviceToDevice, h->s));
  return GLSU_OK;
}

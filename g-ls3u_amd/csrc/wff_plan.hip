// wff_plan.hip -- WFF2 plan set-up (window spectra, twiddles, buffers) and threshold.
#include "wff_kernels.h"

namespace glsu {

#define GLSU_PLAN_INFO(id, T, E, ...) \
  {id, T, E, T * E, #__VA_ARGS__, FftPlan<T, E, __VA_ARGS__>::XL0 ? 1 : 0},
const std::vector<FftPlanInfo>& fft_plans() {
  static const std::vector<FftPlanInfo> v = {GLSU_PLANS(GLSU_PLAN_INFO)};
  return v;
}

std::vector<int> fft_candidates(int need) {
  std::vector<std::pair<int, int>> c;
  for (auto& p : fft_plans())
    if (p.N >= need) c.push_back({p.N, p.id});
  std::sort(c.begin(), c.end());
  std::vector<int> out;
  for (auto& q : c) out.push_back(q.second);
  return out;
}

// ------------------------------------------------------------ threshold
// thr = 6 * sqrt(sum |f|^2 / (W*H) / 3), WFT/WFT2_CPU.cpp:982-1016.
__global__ __launch_bounds__(256) void thr_partial_kernel(const float2* __restrict__ f, size_t n,
                                                          float* __restrict__ part) {
  float s = 0.f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float2 v = f[i];
    s += v.x * v.x + v.y * v.y;
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void thr_final_kernel(const float* __restrict__ part, int nparts,
                                                        float inv_n, float* __restrict__ thr) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *thr = 6.f * sqrtf(red[0] * inv_n / 3.f);
}

hipError_t launch_thr(const WffPlan& p, const float2* f, hipStream_t s) {
  const size_t n = (size_t)p.W * p.H;
  const int parts = 256;
  hipLaunchKernelGGL(thr_partial_kernel, dim3(parts), dim3(256), 0, s, f, n, p.d_red);
  hipLaunchKernelGGL(thr_final_kernel, dim3(1), dim3(256), 0, s, p.d_red, parts,
                     (float)(1.0 / (double)n), p.d_thr);
  return hipGetLastError();
}

// Batched threshold: blocks [256 s, 256 s + 256) reduce image s exactly as
// thr_partial_kernel does with a 256-block grid (same strides, same tree), so
// each image's threshold is bit-identical to a single-image run.  Images whose
// threshold is already set (>= 0) are skipped.
__global__ __launch_bounds__(256) void thr_partial_batch_kernel(const BatchSeg* __restrict__ tab) {
  const BatchSeg& sg = tab[blockIdx.x / 256];
  if (*sg.thr >= 0.f) return;
  const int b = blockIdx.x % 256;
  const size_t n = (size_t)sg.W * sg.H;
  float s = 0.f;
  for (size_t i = (size_t)b * blockDim.x + threadIdx.x; i < n; i += (size_t)256 * blockDim.x) {
    const float2 v = sg.z[i];
    s += v.x * v.x + v.y * v.y;
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sg.red[b] = red[0];
}

__global__ __launch_bounds__(256) void thr_final_batch_kernel(const BatchSeg* __restrict__ tab) {
  const BatchSeg& sg = tab[blockIdx.x];
  if (*sg.thr >= 0.f) return;
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < 256; i += 256) s += sg.red[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *sg.thr = 6.f * sqrtf(red[0] * sg.inv_n / 3.f);
}

hipError_t launch_thr_batch(const BatchSeg* tab, int nseg, hipStream_t s) {
  hipLaunchKernelGGL(thr_partial_batch_kernel, dim3(256 * nseg), dim3(256), 0, s, tab);
  hipLaunchKernelGGL(thr_final_batch_kernel, dim3(nseg), dim3(256), 0, s, tab);
  return hipGetLastError();
}

__global__ void frame_advance_kernel(int* frame) {
  if (threadIdx.x == 0) *frame += 1;
}

hipError_t launch_frame_advance(int* frame, hipStream_t s) {
  hipLaunchKernelGGL(frame_advance_kernel, dim3(1), dim3(64), 0, s, frame);
  return hipGetLastError();
}

// --------------------------------------------------------- plan set-up
// Smallest compiled plan with N >= need, in (N, id) order (the first of
// fft_candidates(need)); t64_only: plans with one transform per wavefront.
static int smallest_plan(int need, bool t64_only) {
  int best = -1, bestN = 0;
  for (const PlanRow& r : kPlanRows) {
    if (r.N < need || (t64_only && r.T != 64)) continue;
    if (best < 0 || r.N < bestN || (r.N == bestN && r.id < best)) {
      best = r.id;
      bestN = r.N;
    }
  }
  return best;
}

int plan_default(int need, bool t64_only) {
  // prefer one transform per wavefront unless a narrower plan is much shorter
  const int first = smallest_plan(need, t64_only);
  const int best64 = smallest_plan(need, true);
  if (first < 0) return -1;
  if (best64 < 0) return first;
  return (plan_row(first)->N < 0.8 * plan_row(best64)->N) ? first : best64;
}

// Window spectra in the order the kernels read them: entry (r, j, lane) holds
// spectral bin kbase(lane) + koff(j) of row r, so that register j of every
// lane of a transform is one contiguous T-element run (a coalesced load).
template <class P>
static void to_lane_major(std::vector<float>& h, int rows) {
  std::vector<float> out(h.size());
  for (int r = 0; r < rows; ++r)
    for (int j = 0; j < P::E; ++j)
      for (int l = 0; l < P::T; ++l)
        out[(size_t)r * P::N + j * P::T + l] = h[(size_t)r * P::N + P::kbase(l) + P::koff(j)];
  h.swap(out);
}

static bool lane_major(int plan, std::vector<float>& h, int rows) {
  switch (plan) {
#define BODY                \
  to_lane_major<P>(h, rows); \
  return true
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      break;
  }
  return false;
}

// (1/N) * DFT of the centred, modulated 1-D window g1(d) e^{j w d}, d in [-s, s]:
// H[k] = (1/N) sum_d g1(d) e^{j w d} e^{-2 pi i k d / N}
//      = (1/N) sum_d g1(d) cos((w - 2 pi k / N) d)
// g1 is even, so the sine terms of d and -d cancel: the spectrum is real.
static std::vector<float> window_spectrum(const std::vector<double>& g1, int s, double w, int N) {
  std::vector<float> out(N);
  for (int k = 0; k < N; ++k) {
    double re = 0;
    for (int d = -s; d <= s; ++d) {
      long kd = ((long)k * d) % N;
      if (kd < 0) kd += N;
      const double a = w * d - 2.0 * M_PI * (double)kd / (double)N;
      re += g1[d + s] * std::cos(a);
    }
    out[k] = (float)(re / N);
  }
  return out;
}

static std::vector<double> norm_window(double sig, int s) {
  // exp(-d^2 / 2 / sigma^2) scaled to unit discrete L2 norm (WFT/WFT2_CPU.cpp:752-805;
  // the 2-D norm is the product of the 1-D norms)
  std::vector<double> g(2 * s + 1);
  double n2 = 0;
  for (int d = -s; d <= s; ++d) {
    g[d + s] = std::exp(-(double)(d * d) / 2 / sig / sig);
    n2 += g[d + s] * g[d + s];
  }
  const double inv = 1.0 / std::sqrt(n2);
  for (auto& v : g) v *= inv;
  return g;
}

static std::vector<float2> twiddles(int N) {
  std::vector<float2> t(N);
  for (int k = 0; k < N; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)N;
    t[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  return t;
}

// ---------------------------------------------------- tables on the device
// The tables wff_plan_init builds on the host (normalised 1-D windows, window
// spectra in lane-major order, twiddles), computed by kernels in FP64 with the
// host's expressions in the host's order, into memory the caller provides.
// The FP64 exp/cos of the device and of the host C library can differ in the
// last bit, so a table entry can, rarely, round to the neighbouring float.
__global__ void norm_window_kernel(double sig, int s, double* __restrict__ g) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;  // serial sum, host order
  double n2 = 0;
  for (int d = -s; d <= s; ++d) {
    const double v = exp(-(double)(d * d) / 2 / sig / sig);
    g[d + s] = v;
    n2 += v * v;
  }
  const double inv = 1.0 / sqrt(n2);
  for (int i = 0; i < 2 * s + 1; ++i) g[i] *= inv;
}

template <class P>
__global__ __launch_bounds__(256) void spectra_kernel(const double* __restrict__ g, int s,
                                                      double wl, double wi, int rows,
                                                      float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * P::N) return;
  const int r = i / P::N, q = i % P::N;
  const int k = P::kbase(q % P::T) + P::koff(q / P::T);
  const double w = wl + (double)r * wi;
  double re = 0;
  for (int d = -s; d <= s; ++d) {
    long kd = ((long)k * d) % P::N;
    if (kd < 0) kd += P::N;
    const double a = w * d - 2.0 * M_PI * (double)kd / (double)P::N;
    re += g[d + s] * cos(a);
  }
  out[i] = (float)(re / P::N);
}

__global__ __launch_bounds__(256) void twiddle_kernel(int N, float2* __restrict__ t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const double a = 2.0 * M_PI * (double)k / (double)N;
  t[k] = make_float2((float)cos(a), (float)-sin(a));
}

__global__ void set_f32_kernel(float* __restrict__ p, float v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

template <class P>
static void launch_spectra(const double* g, int s, double wl, double wi, int rows, float* out,
                           hipStream_t st) {
  const int n = rows * P::N;
  hipLaunchKernelGGL(spectra_kernel<P>, dim3((n + 255) / 256), dim3(256), 0, st, g, s, wl, wi,
                     rows, out);
}

hipError_t launch_plan_tables(const WffPlan& p, double* gx, double* gy, hipStream_t st) {
  hipLaunchKernelGGL(norm_window_kernel, dim3(1), dim3(64), 0, st, p.sigx, p.sx, gx);
  hipLaunchKernelGGL(norm_window_kernel, dim3(1), dim3(64), 0, st, p.sigy, p.sy, gy);
  switch (p.planx) {
#define BODY launch_spectra<P>(gx, p.sx, p.wxl, p.wxi, p.nwx, p.d_Hx, st)
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  switch (p.plany) {
#define BODY launch_spectra<P>(gy, p.sy, p.wyl, p.wyi, p.nwy, p.d_Hy, st)
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(twiddle_kernel, dim3((p.Nx + 255) / 256), dim3(256), 0, st, p.Nx, p.d_twx);
  hipLaunchKernelGGL(twiddle_kernel, dim3((p.Ny + 255) / 256), dim3(256), 0, st, p.Ny, p.d_twy);
  if (p.thr_param >= 0)
    hipLaunchKernelGGL(set_f32_kernel, dim3(1), dim3(64), 0, st, p.d_thr, (float)p.thr_param);
  return hipGetLastError();
}

template <class V>
static hipError_t upload(V*& dptr, const std::vector<V>& h) {
  hipError_t e = hipMalloc((void**)&dptr, sizeof(V) * h.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(dptr, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice);
}

int wff_plan_geometry(WffPlan& p, int W, int H, double sigx, double wxl, double wxh, double wxi,
                      double sigy, double wyl, double wyh, double wyi, double thr, int planx,
                      int plany, int cols_waves, int rows_waves, const char** err) {
  p = WffPlan{};
  p.W = W;
  p.H = H;
  p.sigx = sigx; p.sigy = sigy;
  p.wxl = wxl; p.wxh = wxh; p.wxi = wxi;
  p.wyl = wyl; p.wyh = wyh; p.wyi = wyi;
  p.thr_param = thr;
  // WFT/WFT2_CPU.cpp:701-702, :222-223
  p.sx = (int)std::round(3 * sigx);
  p.sy = (int)std::round(3 * sigy);
  p.nwx = (int)((wxh - wxl) * (1 / wxi)) + 1;
  p.nwy = (int)((wyh - wyl) * (1 / wyi)) + 1;
  if (p.nwx < 1 || p.nwy < 1) {
    *err = "empty frequency grid";
    return -1;
  }
  p.scale = (float)(0.25 * (1.0 / (M_PI * M_PI)) * wxi * wyi);
  const int needx = W + p.sx, needy = H + p.sy;
  if (planx < 0) {
    planx = smallest_plan(needx, false);
    if (planx < 0) { *err = "no FFT plan long enough along x"; return -2; }
  }
  if (plany < 0) {
    plany = smallest_plan(needy, false);
    if (plany < 0) { *err = "no FFT plan long enough along y"; return -2; }
  }
  const PlanRow* px = plan_row(planx);
  const PlanRow* py = plan_row(plany);
  if (!px || !py || px->N < needx || py->N < needy) {
    *err = "FFT plan too short for image + half window";
    return -2;
  }
  p.planx = planx; p.plany = plany;
  p.Nx = px->N; p.Ny = py->N;
  // rows: one wavefront per transform, up to kRowWavesMax waves per row
  p.rows_waves = 1;
  if (rows_waves >= 1 && rows_waves <= kRowWavesMax) p.rows_waves = std::min(rows_waves, p.nwx);
  if (cols_waves == 1 || cols_waves == 2 || cols_waves == 4) p.cols_waves = cols_waves;
  // keep dynamic LDS per workgroup within 64 KiB
  while (p.cols_waves > 1 && plan_smem_bytes(plany, p.cols_waves, true) > 65536) p.cols_waves /= 2;
  while (p.rows_waves > 1 && plan_smem_bytes(planx, p.rows_waves) > 65536) --p.rows_waves;
  const int wb = (W + kTileCols - 1) / kTileCols;
  p.plane_elems = (size_t)wb * kTileCols * H;
  return 0;
}

int wff_plan_init(WffPlan& p, int W, int H, double sigx, double wxl, double wxh, double wxi,
                  double sigy, double wyl, double wyh, double wyi, double thr, int planx,
                  int plany, int cols_waves, int rows_waves, std::string& err,
                  bool buffers) {
  const char* gerr = nullptr;
  if (int rc = wff_plan_geometry(p, W, H, sigx, wxl, wxh, wxi, sigy, wyl, wyh, wyi, thr, planx,
                                 plany, cols_waves, rows_waves, &gerr)) {
    err = gerr;
    return rc;
  }
  // spectra tables
  auto gx = norm_window(sigx, p.sx), gy = norm_window(sigy, p.sy);
  std::vector<float> hx, hy;
  for (int k = 0; k < p.nwx; ++k) {
    auto t = window_spectrum(gx, p.sx, wxl + (double)k * wxi, p.Nx);
    hx.insert(hx.end(), t.begin(), t.end());
  }
  for (int k = 0; k < p.nwy; ++k) {
    auto t = window_spectrum(gy, p.sy, wyl + (double)k * wyi, p.Ny);
    hy.insert(hy.end(), t.begin(), t.end());
  }
  if (!lane_major(planx, hx, p.nwx) || !lane_major(plany, hy, p.nwy)) {
    err = "unknown FFT plan";
    return -2;
  }
  hipError_t e;
  if ((e = upload(p.d_Hx, hx)) != hipSuccess || (e = upload(p.d_Hy, hy)) != hipSuccess ||
      (e = upload(p.d_twx, twiddles(p.Nx))) != hipSuccess ||
      (e = upload(p.d_twy, twiddles(p.Ny))) != hipSuccess) {
    err = std::string("device allocation failed: ") + hipGetErrorString(e);
    wff_plan_free(p);
    return -4;
  }
  if (!buffers) return 0;
  if ((e = hipMalloc((void**)&p.d_A, sizeof(float2) * p.plane_elems * p.nwx)) != hipSuccess ||
      (e = hipMalloc((void**)&p.d_D, sizeof(float2) * p.plane_elems * p.nwx)) != hipSuccess ||
      (e = hipMalloc((void**)&p.d_thr, sizeof(float))) != hipSuccess ||
      (e = hipMalloc((void**)&p.d_red, sizeof(float) * 256)) != hipSuccess) {
    err = std::string("device allocation failed: ") + hipGetErrorString(e);
    wff_plan_free(p);
    return -4;
  }
  if (thr >= 0) {
    const float t = (float)thr;
    hipMemcpy(p.d_thr, &t, sizeof(float), hipMemcpyHostToDevice);
    p.thr_ready = true;
  }
  return 0;
}

void wff_plan_free(WffPlan& p) {
  hipFree(p.d_Hx); hipFree(p.d_Hy); hipFree(p.d_twx); hipFree(p.d_twy);
  hipFree(p.d_A); hipFree(p.d_D); hipFree(p.d_thr); hipFree(p.d_red);
  p.d_Hx = p.d_Hy = nullptr;
  p.d_twx = p.d_twy = p.d_A = p.d_D = nullptr;
  p.d_thr = p.d_red = nullptr;
}

}  // namespace glsu

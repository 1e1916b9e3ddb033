// example_classes.cpp -- C++ host program using the reference-style classes of
// include/glsu.hpp beyond the DPRA loop, the way reference code calls them:
//   AIA::AIA_HIPF(M, W, H) and operator()(phi, deltas, time, iters, err,
//     frames, maxIter, maxErr) on M phase-shifted 8-bit frames
//     (App_DPRA/dpraworker.cpp:21-57 calls the CPU twin the same way);
//   WFT_FPA::WFT::WFT2_HIPF in WFF mode: the constructor allocates
//     WFT2_DeviceResultsF::m_d_filtered and the struct's destructor frees it
//     (WFT/WFT2_CUDAf.h:21-40, WFT/WFT.cpp:118-130); checked on a plane wave
//     whose frequency is on the (wx, wy) grid, where the filtered phase equals
//     the input phase at pixels 2s or more from the border;
//   WFT_FPA::WFT::WFT2_HIPF in WFR mode with explicit parameters, filling
//     WFT2_DeviceResultsF's ridge maps (WFT/WFT2_CUDAf.h:21-40).
// Prints one line per check with the measured error; exits 0 when all pass.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/glsu.hpp"

static double wrap(double a) { return std::atan2(std::sin(a), std::cos(a)); }

static int run_aia() {
  const int W = 192, H = 160, M = 6;
  const double shifts[M] = {0.0, 1.1, 2.3, -2.6, -1.2, 0.45};
  std::vector<std::vector<std::uint8_t>> frames(M, std::vector<std::uint8_t>((size_t)W * H));
  std::vector<double> phi((size_t)W * H);
  for (int i = 0; i < H; ++i)
    for (int j = 0; j < W; ++j) {
      const double p = 0.0007 * ((j - 90.0) * (j - 90.0) + (i - 70.0) * (i - 70.0)) + 0.25 * j;
      phi[(size_t)i * W + j] = p;
      for (int k = 0; k < M; ++k)
        frames[k][(size_t)i * W + j] = (std::uint8_t)std::lround(128 + 100 * std::cos(p + shifts[k]));
    }
  std::vector<const std::uint8_t*> f;
  for (auto& v : frames) f.push_back(v.data());
  std::vector<float> v_phi, v_deltas = {0.0f, 1.0f, 2.0f, -2.5f, -1.0f, 0.5f};
  double t = 0;
  int iters = 0;
  float err = 0;
  AIA::AIA_HIPF aia(M, W, H);
  aia(v_phi, v_deltas, t, iters, err, f, 20, 1e-4f, 1, /*err_mode*/ 1);
  double dmax = 0, prms = 0;
  for (int k = 0; k < M; ++k) dmax = std::fmax(dmax, std::fabs(wrap(v_deltas[k] - shifts[k])));
  for (size_t q = 0; q < phi.size(); ++q) {
    const double e = wrap(v_phi[q] - (phi[q] + shifts[0]));
    prms += e * e;
  }
  prms = std::sqrt(prms / phi.size());
  std::printf("aia: %d iterations, err %.2e, %.3f ms; shift error %.2e, phase rms %.2e\n", iters,
              err, t, dmax, prms);
  return (dmax < 0.02 && prms < 0.02) ? 0 : 1;
}

static int run_wff() {
  const int W = 96, H = 80, s = 12;  // sigma 4 -> half window round(3 sigma) = 12
  const double wx0 = 0.1, wy0 = -0.2;
  std::vector<float2> f((size_t)W * H);
  for (int i = 0; i < H; ++i)
    for (int j = 0; j < W; ++j) {
      const double p = wx0 * j + wy0 * i - 0.7;
      f[(size_t)i * W + j] = make_float2((float)std::cos(p), (float)std::sin(p));
    }
  float2* d_f = nullptr;
  if (hipMalloc((void**)&d_f, sizeof(float2) * f.size()) != hipSuccess) return 1;
  (void)hipMemcpy(d_f, f.data(), sizeof(float2) * f.size(), hipMemcpyHostToDevice);
  float2* filtered = nullptr;
  int attr_rc = -1, mem_type = -1;
  std::vector<float2> F(f.size());
  double t = 0;
  {
    WFT_FPA::WFT::WFT2_DeviceResultsF z;
    WFT_FPA::WFT::WFT2_HIPF wff(W, H, WFT_FPA::WFT::WFT_TYPE::WFF, 4.0, -0.2, 0.2, 0.1, 4.0,
                                -0.2, 0.2, 0.1, -1, z);
    filtered = z.m_d_filtered;  // allocated by the constructor
    hipPointerAttribute_t a;
    attr_rc = (int)hipPointerGetAttributes(&a, filtered);
    mem_type = (int)a.type;
    wff(d_f, z, t);
    (void)hipMemcpy(F.data(), z.m_d_filtered, sizeof(float2) * F.size(), hipMemcpyDeviceToHost);
  }  // ~WFT2_HIPF, then ~WFT2_DeviceResultsF frees m_d_filtered
  (void)hipFree(d_f);
  double ep = 0, amin = 1e30;
  for (int i = 2 * s; i < H - 2 * s; ++i)
    for (int j = 2 * s; j < W - 2 * s; ++j) {
      const size_t q = (size_t)i * W + j;
      const double ph = std::atan2((double)F[q].y, (double)F[q].x);
      ep = std::fmax(ep, std::fabs(wrap(ph - (wx0 * j + wy0 * i - 0.7))));
      amin = std::fmin(amin, std::hypot((double)F[q].x, (double)F[q].y));
    }
  const bool allocated = filtered != nullptr && attr_rc == 0 && mem_type == (int)hipMemoryTypeDevice;
  std::printf("wff: %.3f ms; filtered field allocated by the constructor: %s; interior phase "
              "error %.2e, min |F| %.3f\n", t, allocated ? "yes" : "no", ep, amin);
  return (allocated && ep < 1e-5 && amin > 0.1) ? 0 : 1;
}

static int run_wfr() {
  const int W = 96, H = 80;
  const double wx0 = 0.2, wy0 = -0.1;
  std::vector<float2> f((size_t)W * H);
  for (int i = 0; i < H; ++i)
    for (int j = 0; j < W; ++j) {
      const double p = wx0 * j + wy0 * i + 0.4;
      f[(size_t)i * W + j] = make_float2((float)std::cos(p), (float)std::sin(p));
    }
  float2* d_f = nullptr;
  if (hipMalloc((void**)&d_f, sizeof(float2) * f.size()) != hipSuccess) return 1;
  (void)hipMemcpy(d_f, f.data(), sizeof(float2) * f.size(), hipMemcpyHostToDevice);
  WFT_FPA::WFT::WFT2_DeviceResultsF z;
  WFT_FPA::WFT::WFT2_HIPF wfr(W, H, WFT_FPA::WFT::WFT_TYPE::WFR, 4.0, -0.3, 0.3, 0.1, 4.0, -0.3,
                              0.3, 0.1, -1, z);
  double t = 0;
  wfr(d_f, z, t);
  std::vector<float> wx(f.size()), wy(f.size()), ph(f.size()), cxx(f.size());
  (void)hipMemcpy(wx.data(), z.m_d_wx, sizeof(float) * f.size(), hipMemcpyDeviceToHost);
  (void)hipMemcpy(wy.data(), z.m_d_wy, sizeof(float) * f.size(), hipMemcpyDeviceToHost);
  (void)hipMemcpy(ph.data(), z.m_d_phase, sizeof(float) * f.size(), hipMemcpyDeviceToHost);
  (void)hipMemcpy(cxx.data(), z.m_d_cxx, sizeof(float) * f.size(), hipMemcpyDeviceToHost);
  (void)hipFree(d_f);
  double ef = 0, ep = 0, ec = 0;
  for (int i = 24; i < H - 24; ++i)
    for (int j = 24; j < W - 24; ++j) {
      const size_t q = (size_t)i * W + j;
      ef = std::fmax(ef, std::fabs(wx[q] - wx0) + std::fabs(wy[q] - wy0));
      ep = std::fmax(ep, std::fabs(wrap(ph[q] - (wx0 * j + wy0 * i + 0.4))));
      ec = std::fmax(ec, std::fabs(cxx[q]));
    }
  std::printf("wfr: %.3f ms; frequency error %.2e, phase error %.2e, curvature %.2e\n", t, ef, ep,
              ec);
  return (ef < 1e-6 && ep < 1e-4 && ec < 1e-4) ? 0 : 1;
}

int main() {
  try {
    const int a = run_aia();
    const int f = run_wff();
    const int w = run_wfr();
    std::printf("example_classes: %s\n", (a | f | w) ? "FAILED" : "ok");
    return a | f | w;
  } catch (int e) {
    std::printf("example_classes: constructor threw %d\n", e);
    return 2;
  } catch (const std::exception& e) {
    std::printf("example_classes: %s\n", e.what());
    return 1;
  }
}

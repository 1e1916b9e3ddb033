// glsu.hpp -- C++ host classes over the C ABI (glsu.h), mirroring the
// reference's class interfaces for the G-LS3U hot path:
//   WFT_FPA::WFT::WFT2_CUDAF  (WFT/WFT2_CUDAf.h:11-40)   -> WFT_FPA::WFT::WFT2_HIPF
//   WFT_FPA::WFT::WFT2_DeviceResultsF (WFT/WFT.h:114-135) -> same name (WFF member)
//   DPRA::DPRA_CUDAF          (DPRA/dpra_cudaf.h:21-45)   -> DPRA::DPRA_HIPF
//   DPRA::DPRA_HYBRIDF        (DPRA/dpra_hybridf.h:14-40) -> DPRA::DPRA_HYBRID_HIPF
//   AIA::AIA_CUDAF            (AIA/aia_cudaf.h:15-72)     -> AIA::AIA_HIPF
// Error behaviour follows the reference: constructors throw int -1 when the
// padded FFT size is out of range (WFT/WFT2_CPU.cpp:69-73); other failures,
// which the reference prints or ignores, throw std::runtime_error.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "glsu.h"

namespace glsu_detail {
inline void check(int rc) {
  if (rc == GLSU_OK) return;
  if (rc == GLSU_ERR_PADDING) throw -1;  // as the reference constructors
  throw std::runtime_error(std::string("glsu: ") + glsu_last_error());
}
}  // namespace glsu_detail

namespace WFT_FPA {
namespace WFT {

enum class WFT_TYPE { WFF, WFR };  // Utils/WFT-FPA.h:72-76

// Device results.  As in the reference, the WFT2 constructor allocates the
// members of its mode -- m_d_filtered (height*width complex64) in WFF mode,
// the eight height*width float maps in WFR mode (WFT/WFT.h:114-135) -- and
// this struct frees them.
struct WFT2_DeviceResultsF {
  float2* m_d_filtered = nullptr;
  float *m_d_wx = nullptr, *m_d_wy = nullptr, *m_d_phase = nullptr, *m_d_phase_comp = nullptr,
        *m_d_b = nullptr, *m_d_r = nullptr, *m_d_cxx = nullptr, *m_d_cyy = nullptr;
  WFT2_DeviceResultsF() = default;
  WFT2_DeviceResultsF(const WFT2_DeviceResultsF&) = delete;
  WFT2_DeviceResultsF& operator=(const WFT2_DeviceResultsF&) = delete;
  ~WFT2_DeviceResultsF() {
    if (m_d_filtered) (void)hipFree(m_d_filtered);
    for (float* q : {m_d_wx, m_d_wy, m_d_phase, m_d_phase_comp, m_d_b, m_d_r, m_d_cxx, m_d_cyy})
      if (q) (void)hipFree(q);
  }
};

class WFT2_HIPF {
 public:
  WFT2_HIPF() = delete;
  WFT2_HIPF(const WFT2_HIPF&) = delete;
  WFT2_HIPF& operator=(const WFT2_HIPF&) = delete;

  // Parameters as in WFT2_CUDAF(W, H, type, sx, wxl, wxh, wxi, sy, wyl, wyh, wyi,
  // thr, z, nStreams); the library runs on one stream (`stream`, default null).
  WFT2_HIPF(int iWidth, int iHeight, WFT_TYPE type, double rSigmaX, double rWxl, double rWxh,
            double rWxi, double rSigmaY, double rWyl, double rWyh, double rWyi, double rThr,
            WFT2_DeviceResultsF& z, int /*iNumStreams*/ = 1, hipStream_t stream = nullptr)
      : w_(iWidth), h_(iHeight), type_(type) {
    glsu_wff_params p{rSigmaX, rWxl, rWxh, rWxi, rSigmaY, rWyl, rWyh, rWyi, rThr};
    glsu_detail::check(glsu_wft2_create(&h, iWidth, iHeight,
                                        type == WFT_TYPE::WFF ? GLSU_WFF : GLSU_WFR, &p, stream));
    alloc(z);
  }
  // Defaults of the type's constructor (WFT/WFT2_CPU.cpp:43-67).
  WFT2_HIPF(int iWidth, int iHeight, WFT_TYPE type, WFT2_DeviceResultsF& z, int = 1,
            hipStream_t stream = nullptr)
      : w_(iWidth), h_(iHeight), type_(type) {
    glsu_wff_params p;
    if (type == WFT_TYPE::WFF)
      glsu_wff_params_default(&p);
    else
      glsu_wfr_params_default(&p);
    glsu_detail::check(glsu_wft2_create(&h, iWidth, iHeight,
                                        type == WFT_TYPE::WFF ? GLSU_WFF : GLSU_WFR, &p, stream));
    alloc(z);
  }
  ~WFT2_HIPF() { glsu_wft2_destroy(h); }

  // operator()(d_f, d_z, time), WFT/WFT2_CUDAf.h:38-40; d_f: device complex64.
  // time: milliseconds of device time (the call synchronises, as the reference).
  void operator()(const float2* d_f, WFT2_DeviceResultsF& d_z, double& time) {
    if (type_ == WFT_TYPE::WFF) {
      glsu_detail::check(glsu_wft2_compute(h, d_f, d_z.m_d_filtered, &time));
    } else {
      const glsu_wfr_results r{d_z.m_d_wx, d_z.m_d_wy, d_z.m_d_phase, d_z.m_d_phase_comp,
                               d_z.m_d_b,  d_z.m_d_r,  d_z.m_d_cxx,   d_z.m_d_cyy};
      glsu_detail::check(glsu_wft2_compute_wfr(h, d_f, &r, &time));
    }
  }

 private:
  void alloc(WFT2_DeviceResultsF& z) {
    const size_t n = (size_t)w_ * h_;
    if (type_ == WFT_TYPE::WFF) {
      if (!z.m_d_filtered && hipMalloc((void**)&z.m_d_filtered, sizeof(float2) * n) != hipSuccess)
        throw std::runtime_error("glsu: allocation of the filtered field failed");
      return;
    }
    for (float** q : {&z.m_d_wx, &z.m_d_wy, &z.m_d_phase, &z.m_d_phase_comp, &z.m_d_b, &z.m_d_r,
                      &z.m_d_cxx, &z.m_d_cyy})
      if (!*q && hipMalloc((void**)q, sizeof(float) * n) != hipSuccess)
        throw std::runtime_error("glsu: allocation of the ridge maps failed");
  }
  glsu_wft2* h = nullptr;
  int w_, h_;
  WFT_TYPE type_;
};

}  // namespace WFT
}  // namespace WFT_FPA

namespace DPRA {

class DPRA_HIPF {
 public:
  DPRA_HIPF() = delete;
  DPRA_HIPF(const DPRA_HIPF&) = delete;
  DPRA_HIPF& operator=(const DPRA_HIPF&) = delete;

  // DPRA_CUDAF(Phi0 (host), W, H, rr): the reference's hard-coded WFF parameters.
  DPRA_HIPF(const float* v_Phi0, int iWidth, int iHeight, int irefUpdateRate = 1,
            hipStream_t stream = nullptr)
      : w_(iWidth), h_(iHeight), rr_(irefUpdateRate) {
    glsu_detail::check(
        glsu_dpra_create_f32(&h, v_Phi0, iWidth, iHeight, irefUpdateRate, nullptr, stream));
  }
  // Same with explicit WFF parameters and an FP64 initial phase.
  DPRA_HIPF(const double* v_Phi0, int iWidth, int iHeight, int irefUpdateRate,
            const glsu_wff_params& p, hipStream_t stream = nullptr)
      : w_(iWidth), h_(iHeight), rr_(irefUpdateRate) {
    glsu_detail::check(glsu_dpra_create(&h, v_Phi0, iWidth, iHeight, irefUpdateRate, &p, stream));
  }
  ~DPRA_HIPF() { glsu_dpra_destroy(h); }

  // dpra_per_frame(img, dPhi, time), DPRA/dpra_cudaf.h:41-43: host frame in,
  // host dPhi out (PCIe-inclusive, synchronising), time in milliseconds.
  void dpra_per_frame(const std::uint8_t* img, std::vector<float>& dPhi, double& time) {
    per_frame(img, GLSU_U8, dPhi, time);
  }
  void dpra_per_frame(const float* img, std::vector<float>& dPhi, double& time) {
    per_frame(img, GLSU_F32, dPhi, time);
  }
  void dpra_per_frame(const double* img, std::vector<float>& dPhi, double& time) {
    per_frame(img, GLSU_F64, dPhi, time);
  }
  void update_ref_phi() { glsu_detail::check(glsu_dpra_update_ref(h)); }

  // operator()(frames, dPhi_Sum, time) with the reference's i % rr updates
  // (DPRA/dpra_cpu.cpp:58-89); frames are host pointers of width*height bytes.
  void operator()(const std::vector<const std::uint8_t*>& f,
                  std::vector<std::vector<float>>& dPhi_Sum, double& time) {
    time = 0;
    std::vector<float> d;
    for (size_t i = 0; i < f.size(); ++i) {
      double t = 0;
      dpra_per_frame(f[i], d, t);
      time += t;
      dPhi_Sum.push_back(d);
      if (i % rr_ == 0) update_ref_phi();
    }
  }

  glsu_dpra* handle() const { return h; }

 private:
  void per_frame(const void* img, int dt, std::vector<float>& dPhi, double& time) {
    dPhi.resize((size_t)w_ * h_);
    glsu_detail::check(glsu_dpra_frame_host(h, img, dt, dPhi.data(), &time));
  }
  glsu_dpra* h = nullptr;
  int w_, h_;
  int rr_;
};

// DPRA::DPRA_HYBRIDF (DPRA/dpra_hybridf.h:14-40, dpra_hybridf.cpp:182-299)
// builds the 3x3 systems on the GPU, copies them to the host, solves them with
// LAPACKE_ssysv on iNumThreads OpenMP threads and copies the field back for
// WFF2 on the GPU.  Same algorithm and parameters as DPRA_CUDAF; here every
// step stays on the device (the host round trip per frame buys nothing on this
// part), so the class is the GPU path under the hybrid's constructor.
class DPRA_HYBRID_HIPF : public DPRA_HIPF {
 public:
  DPRA_HYBRID_HIPF(const float* v_Phi0, int iWidth, int iHeight, int irefUpdateRate = 1,
                   int /*iNumThreads*/ = 1, hipStream_t stream = nullptr)
      : DPRA_HIPF(v_Phi0, iWidth, iHeight, irefUpdateRate, stream) {}
};

}  // namespace DPRA

namespace glsu_output {
// Phase to 8-bit video frames on the device (App_DPRA/dprawidget.cpp:541-549):
// n pixels of d_phi (float) to d_out (uint8), asynchronous on `stream`.
inline void phase_to_u8(const float* d_phi, std::uint8_t* d_out, size_t n,
                        hipStream_t stream = nullptr) {
  glsu_detail::check(glsu_phase_to_u8(d_phi, d_out, n, stream));
}
}  // namespace glsu_output

namespace AIA {

// AIA::AIA_CUDAF (AIA/aia_cudaf.h:15-72) -> AIA::AIA_HIPF: the initial phase
// and the phase shifts of iM >= 3 randomly phase-shifted frames.
class AIA_HIPF {
 public:
  AIA_HIPF() = delete;
  AIA_HIPF(const AIA_HIPF&) = delete;
  AIA_HIPF& operator=(const AIA_HIPF&) = delete;

  // AIA_CUDAF(iM, icols, irows), AIA/aia_cudaf.h:23-25.
  AIA_HIPF(int iM, int icols, int irows, hipStream_t stream = nullptr)
      : m_(iM), w_(icols), h_(irows) {
    glsu_detail::check(glsu_aia_create(&h, iM, icols, irows, stream));
  }
  ~AIA_HIPF() { glsu_aia_destroy(h); }

  // operator()(v_phi, v_deltas, runningtime, iters, err, v_f, iMaxIterations,
  // dMaxErr, iNumThreads), AIA/aia_cudaf.h:28-38: host frames of icols*irows
  // bytes (the reference's cv::Mat data); v_deltas holds the initial shifts when
  // its size is iM, else they are drawn (delta_0 = 0, N(0,1) from `seed`; the
  // reference seeds from the clock).  err_mode 0 is AIA_CUDAF's error metric,
  // 1 that of AIA_CPU_DnF.  A normal matrix that is not positive definite
  // throws std::runtime_error (the reference prints and exits).
  void operator()(std::vector<float>& v_phi, std::vector<float>& v_deltas, double& runningtime,
                  int& iters, float& err, const std::vector<const std::uint8_t*>& v_f,
                  int iMaxIterations = 20, float dMaxErr = 1e-4f, int /*iNumThreads*/ = 1,
                  int err_mode = 0, unsigned long long seed = 0) {
    if ((int)v_f.size() != m_) throw std::runtime_error("glsu: AIA frame count differs");
    const int given = (int)v_deltas.size() == m_;
    v_deltas.resize(m_);
    v_phi.resize((size_t)w_ * h_);
    std::vector<const void*> f(v_f.begin(), v_f.end());
    double e = 0;
    glsu_detail::check(glsu_aia_run_host(h, f.data(), GLSU_U8, v_deltas.data(), given, seed,
                                         iMaxIterations, dMaxErr, err_mode, v_phi.data(), &iters,
                                         &e, &runningtime));
    err = (float)e;
  }

  glsu_aia* handle() const { return h; }

 private:
  glsu_aia* h = nullptr;
  int m_, w_, h_;
};

}  // namespace AIA

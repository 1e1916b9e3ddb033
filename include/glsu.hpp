// glsu.hpp -- C++ host classes over the C ABI (glsu.h), mirroring the
// reference's class interfaces for the G-LS3U hot path:
//   WFT_FPA::WFT::WFT2_CUDAF  (WFT/WFT2_CUDAf.h:11-40)   -> WFT_FPA::WFT::WFT2_HIPF
//   WFT_FPA::WFT::WFT2_DeviceResultsF (WFT/WFT.h:114-135) -> same name (WFF member)
//   DPRA::DPRA_CUDAF          (DPRA/dpra_cudaf.h:21-45)   -> DPRA::DPRA_HIPF
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

// Device results of the WFF mode.  As in the reference, the WFT2 constructor
// allocates m_d_filtered (height*width complex64) and this struct frees it.
struct WFT2_DeviceResultsF {
  float2* m_d_filtered = nullptr;
  WFT2_DeviceResultsF() = default;
  WFT2_DeviceResultsF(const WFT2_DeviceResultsF&) = delete;
  WFT2_DeviceResultsF& operator=(const WFT2_DeviceResultsF&) = delete;
  ~WFT2_DeviceResultsF() {
    if (m_d_filtered) (void)hipFree(m_d_filtered);
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
      : w_(iWidth), h_(iHeight) {
    glsu_wff_params p{rSigmaX, rWxl, rWxh, rWxi, rSigmaY, rWyl, rWyh, rWyi, rThr};
    glsu_detail::check(glsu_wft2_create(&h, iWidth, iHeight,
                                        type == WFT_TYPE::WFF ? GLSU_WFF : GLSU_WFR, &p, stream));
    alloc(z);
  }
  // Defaults of the WFF-mode constructor (WFT/WFT2_CPU.cpp:43-54).
  WFT2_HIPF(int iWidth, int iHeight, WFT_TYPE type, WFT2_DeviceResultsF& z, int = 1,
            hipStream_t stream = nullptr)
      : w_(iWidth), h_(iHeight) {
    glsu_wff_params p;
    glsu_wff_params_default(&p);
    glsu_detail::check(glsu_wft2_create(&h, iWidth, iHeight,
                                        type == WFT_TYPE::WFF ? GLSU_WFF : GLSU_WFR, &p, stream));
    alloc(z);
  }
  ~WFT2_HIPF() { glsu_wft2_destroy(h); }

  // operator()(d_f, d_z, time), WFT/WFT2_CUDAf.h:38-40; d_f: device complex64.
  // time: milliseconds of device time (the call synchronises, as the reference).
  void operator()(const float2* d_f, WFT2_DeviceResultsF& d_z, double& time) {
    glsu_detail::check(glsu_wft2_compute(h, d_f, d_z.m_d_filtered, &time));
  }

 private:
  void alloc(WFT2_DeviceResultsF& z) {
    if (!z.m_d_filtered &&
        hipMalloc((void**)&z.m_d_filtered, sizeof(float2) * (size_t)w_ * h_) != hipSuccess)
      throw std::runtime_error("glsu: allocation of the filtered field failed");
  }
  glsu_wft2* h = nullptr;
  int w_, h_;
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

}  // namespace DPRA

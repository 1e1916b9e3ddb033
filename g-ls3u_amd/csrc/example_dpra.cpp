// example_dpra.cpp -- C++ host program using the reference-style classes of
// include/glsu.hpp: the DPRA + WFF2 chain over synthetic 8-bit fringe frames,
// as the reference demo's frame loop does (App_DPRA/dprawidget.cpp:241-297:
// construct once, dpra_per_frame per frame, update_ref_phi when i % rr == 0).
// Usage: example_dpra [width height frames [hybrid]]
// "hybrid" runs the loop through DPRA::DPRA_HYBRID_HIPF (the DPRA_HYBRIDF
// constructor: reference WFF parameters); the last frame's cumulative phase is
// also turned into an 8-bit video frame on the device.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/glsu.hpp"

static double peaks(double X, double Y) {
  return 3.0 * (1 - X) * (1 - X) * std::exp(-X * X - (Y + 1) * (Y + 1)) -
         10.0 * (X / 5 - X * X * X - std::pow(Y, 5)) * std::exp(-X * X - Y * Y) -
         std::exp(-(X + 1) * (X + 1) - Y * Y) / 3.0;
}

int main(int argc, char** argv) {
  const int W = argc > 1 ? std::atoi(argv[1]) : 256;
  const int H = argc > 2 ? std::atoi(argv[2]) : 256;
  const int T = argc > 3 ? std::atoi(argv[3]) : 5;
  const bool hybrid = argc > 4 && std::string(argv[4]) == "hybrid";
  std::vector<float> phi0((size_t)W * H);
  std::vector<std::vector<std::uint8_t>> frames(T, std::vector<std::uint8_t>((size_t)W * H));
  for (int t = 0; t < T; ++t)
    for (int i = 0; i < H; ++i)
      for (int j = 0; j < W; ++j) {
        const double X = -3.0 + 6.0 * j / (W - 1), Y = -3.0 + 6.0 * i / (H - 1);
        const double ph = (3.0 + 0.1 * t) * peaks(X, Y) + 2 * M_PI * 0.05 * j;
        if (t == 0) phi0[(size_t)i * W + j] = (float)ph;
        frames[t][(size_t)i * W + j] = (std::uint8_t)std::lround(128 + 100 * std::cos(ph));
      }
  try {
    glsu_wff_params p{10, -0.2, 0.2, 0.1, 10, -0.2, 0.2, 0.1, -1};
    std::vector<double> phi0d(phi0.begin(), phi0.end());
    std::vector<const std::uint8_t*> ptrs;
    for (auto& f : frames) ptrs.push_back(f.data());
    std::vector<std::vector<float>> dphi;
    double ms = 0;
    if (hybrid) {
      DPRA::DPRA_HYBRID_HIPF dpra(phi0.data(), W, H, 1, /*iNumThreads*/ 8);
      dpra(ptrs, dphi, ms);
    } else {
      DPRA::DPRA_HIPF dpra(phi0d.data(), W, H, 1, p);
      dpra(ptrs, dphi, ms);
    }
    // output stage on the device: the last frame as an 8-bit video frame
    float* d_phi = nullptr;
    std::uint8_t* d_u8 = nullptr;
    std::vector<std::uint8_t> frame((size_t)W * H);
    if (hipMalloc((void**)&d_phi, sizeof(float) * frame.size()) != hipSuccess ||
        hipMalloc((void**)&d_u8, frame.size()) != hipSuccess)
      throw std::runtime_error("hipMalloc failed");
    (void)hipMemcpy(d_phi, dphi.back().data(), sizeof(float) * frame.size(), hipMemcpyHostToDevice);
    glsu_output::phase_to_u8(d_phi, d_u8, frame.size());
    (void)hipMemcpy(frame.data(), d_u8, frame.size(), hipMemcpyDeviceToHost);
    (void)hipFree(d_phi);
    (void)hipFree(d_u8);
    unsigned long long sum = 0;
    for (auto v : frame) sum += v;
    // phase change of the last frame at the centre vs its closed form 0.1*(T-1)*peaks
    const int c = (H / 2) * W + W / 2;
    const double X = -3.0 + 6.0 * (W / 2) / (W - 1), Y = -3.0 + 6.0 * (H / 2) / (H - 1);
    std::printf("example_dpra: %s %dx%d, %d frames, %.3f ms; centre dphi %.4f (model %.4f); "
                "video frame mean %.2f\n",
                hybrid ? "hybrid" : "plain", W, H, T, ms, dphi.back()[c],
                0.1 * (T - 1) * peaks(X, Y), (double)sum / frame.size());
  } catch (int e) {
    std::printf("example_dpra: constructor threw %d (padding out of range)\n", e);
    return 2;
  } catch (const std::exception& e) {
    std::printf("example_dpra: %s\n", e.what());
    return 1;
  }
  return 0;
}

// Host check of the in-register DFTs of fft_device.h (Dft<N, INV>) against a
// direct FP64 DFT: prime sizes, prime-factor (Good-Thomas) composites and
// Cooley-Tukey composites, both directions.  Exit status 0 when every size is
// within 2e-6 of the FP64 result, relative to the largest output.
#include "fft_device.h"

#include <cmath>
#include <cstdio>
#include <random>

using glsu::Dft;

template <int N, bool INV>
static double check(std::mt19937& g) {
  std::normal_distribution<float> d;
  float2 x[N], y[N];
  for (int i = 0; i < N; ++i) x[i] = y[i] = make_float2(d(g), d(g));
  Dft<N, INV>::run(y);
  double err = 0, nrm = 0;
  for (int k = 0; k < N; ++k) {
    double re = 0, im = 0;
    for (int n = 0; n < N; ++n) {
      const double a = (INV ? 2 : -2) * M_PI * (double)((long)n * k % N) / N;
      re += x[n].x * std::cos(a) - x[n].y * std::sin(a);
      im += x[n].x * std::sin(a) + x[n].y * std::cos(a);
    }
    err = std::fmax(err, std::hypot(y[k].x - re, y[k].y - im));
    nrm = std::fmax(nrm, std::hypot(re, im));
  }
  return err / nrm;
}

int main() {
  std::mt19937 g(1);
  double worst = 0;
#define GLSU_CHECK(n)                                                            \
  {                                                                              \
    const double e = std::fmax(check<n, false>(g), check<n, true>(g));           \
    std::printf("N=%d rel=%.3e\n", n, e);                                       \
    worst = std::fmax(worst, e);                                                 \
  }
  GLSU_CHECK(2) GLSU_CHECK(3) GLSU_CHECK(4) GLSU_CHECK(5) GLSU_CHECK(6) GLSU_CHECK(8)
  GLSU_CHECK(9) GLSU_CHECK(10) GLSU_CHECK(12) GLSU_CHECK(15) GLSU_CHECK(16) GLSU_CHECK(20)
  GLSU_CHECK(24) GLSU_CHECK(28) GLSU_CHECK(32) GLSU_CHECK(36) GLSU_CHECK(40)
  std::printf("worst %.3e\n", worst);
  return worst < 2e-6 ? 0 : 1;
}

// fft_device.h -- wave-level complex FFTs for gfx950 (CDNA4), FP32.
//
// A transform of length N = T * E is owned by T lanes of one wavefront; lane
// t keeps the E elements x[t + T*j] (j < E) in VGPRs.  The first Stockham
// stage is an E-point DFT done entirely in registers (compile-time twiddles);
// the remaining stages are radix-R butterflies (R | E, product of the R's = T)
// that exchange through a per-transform LDS buffer with one padding element
// every E elements (keeps the stride-E scatter of stage 1 free of bank
// conflicts).  Every stage reads the natural "t + T*j" pattern, so the last
// stage leaves X[t + T*j] in register j: spectra and signals are always held
// in the same lane/register layout, and pointwise work between transforms
// (spectrum multiplies, hard threshold) needs no data movement.
//
// Stockham indexing (radix r, L = product of earlier radices, b = butterfly):
//   in  = x[b + m*N/r],  twiddle W_{L*r}^{(b mod L)*m},
//   out = y[(b div L)*L*r + (b mod L) + m*L].
// Published algorithm: T. G. Stockham, "High-speed convolution and
// correlation" (1966) autosort formulation.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace glsu {

// ---------------------------------------------------------------- complex
struct cf { float x, y; };

__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// ------------------------------------------------- compile-time trigonometry
// Taylor series after reduction to [-pi, pi]; used only in constant
// expressions (twiddles of the in-register DFTs), evaluated in double.
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double ct_reduce(double a) {
  while (a > kPi) a -= 2 * kPi;
  while (a < -kPi) a += 2 * kPi;
  return a;
}
constexpr double ct_sin(double a) {
  a = ct_reduce(a);
  double term = a, sum = a;
  for (int k = 1; k < 30; ++k) {
    term *= -a * a / ((2.0 * k) * (2.0 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos(double a) {
  a = ct_reduce(a);
  double term = 1, sum = 1;
  for (int k = 1; k < 30; ++k) {
    term *= -a * a / ((2.0 * k - 1) * (2.0 * k));
    sum += term;
  }
  return sum;
}

template <int N, int K, bool INV>
struct Tw {
  static constexpr int k = ((K % N) + N) % N;
  static constexpr double ang = 2.0 * kPi * k / N;
  static constexpr float c = (float)ct_cos(ang);
  static constexpr float s = INV ? (float)ct_sin(ang) : -(float)ct_sin(ang);
};

// multiply by W_N^K (forward sign -, inverse +), special-casing the trivial
// eighth-turn values so they cost no multiplies
template <int N, int K, bool INV>
__host__ __device__ __forceinline__ float2 twc(float2 a) {
  constexpr int k = ((K % N) + N) % N;
  if constexpr (k == 0) {
    return a;
  } else if constexpr ((4 * k) % N == 0) {
    constexpr int q = 4 * k / N;  // quarter turns
    if constexpr (q == 2) return make_float2(-a.x, -a.y);
    else if constexpr ((q == 1) != INV) return make_float2(a.y, -a.x);   // * (-i)
    else return make_float2(-a.y, a.x);                                   // * (+i)
  } else {
    return make_float2(a.x * Tw<N, K, INV>::c - a.y * Tw<N, K, INV>::s,
                       a.x * Tw<N, K, INV>::s + a.y * Tw<N, K, INV>::c);
  }
}

// ------------------------------------------------------ static loop helper
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

constexpr int pick_factor(int n) {
  if (n % 4 == 0 && n != 4) return 4;
  if (n % 2 == 0 && n != 2) return 2;
  for (int p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

// --------------------------------------------- in-register DFT of size N
template <int N, bool INV>
struct Dft;

template <bool INV>
struct Dft<1, INV> {
  __host__ __device__ __forceinline__ static void run(float2*) {}
};
template <bool INV>
struct Dft<2, INV> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};
template <bool INV>
struct Dft<3, INV> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    constexpr float c1 = -0.5f;
    constexpr float s1 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
    float2 a = x[0], b = x[1], c = x[2];
    float2 t = cadd(b, c), d = csub(b, c);
    float2 m = make_float2(a.x + c1 * t.x, a.y + c1 * t.y);
    float2 r = make_float2(-s1 * d.y, s1 * d.x);  // s1 * i * d
    x[0] = cadd(a, t);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};
template <bool INV>
struct Dft<4, INV> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
    float2 t2 = cadd(x[1], x[3]), t3 = csub(x[1], x[3]);
    // t3 * (-i) forward, * (+i) inverse
    float2 t3m = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3m);
    x[3] = csub(t1, t3m);
  }
};
template <bool INV>
struct Dft<5, INV> {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    constexpr float c1 = (float)ct_cos(2 * kPi / 5), c2 = (float)ct_cos(4 * kPi / 5);
    constexpr float s1 = (INV ? 1 : -1) * (float)ct_sin(2 * kPi / 5);
    constexpr float s2 = (INV ? 1 : -1) * (float)ct_sin(4 * kPi / 5);
    float2 a = x[0];
    float2 p1 = cadd(x[1], x[4]), m1 = csub(x[1], x[4]);
    float2 p2 = cadd(x[2], x[3]), m2 = csub(x[2], x[3]);
    float2 r1 = make_float2(a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y);
    float2 r2 = make_float2(a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y);
    // i * (s1*m1 + s2*m2) and i * (s2*m1 - s1*m2)
    float2 q1 = make_float2(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y);
    float2 q2 = make_float2(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y);
    x[0] = make_float2(a.x + p1.x + p2.x, a.y + p1.y + p2.y);
    x[1] = make_float2(r1.x - q1.y, r1.y + q1.x);
    x[4] = make_float2(r1.x + q1.y, r1.y - q1.x);
    x[2] = make_float2(r2.x - q2.y, r2.y + q2.x);
    x[3] = make_float2(r2.x + q2.y, r2.y - q2.x);
  }
};

// generic: prime sizes by direct O(N^2) sums, composite N = P*Q by one
// decimation step X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_N^{n2 k1} sum_n1 x[n2 + Q n1] W_P^{n1 k1}
template <int N, bool INV>
struct Dft {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    constexpr int P = pick_factor(N);
    constexpr int Q = N / P;
    if constexpr (Q == 1) {
      float2 y[N];
      sfor<0, N>([&](auto k) {
        constexpr int K = decltype(k)::value;
        float2 s = x[0];
        sfor<1, N>([&](auto n) {
          constexpr int M = decltype(n)::value;
          s = cadd(s, twc<N, (M * K) % N, INV>(x[M]));
        });
        y[K] = s;
      });
      sfor<0, N>([&](auto k) { x[decltype(k)::value] = y[decltype(k)::value]; });
    } else {
      float2 y[P][Q];
      sfor<0, Q>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        float2 t[P];
        sfor<0, P>([&](auto n1) { t[decltype(n1)::value] = x[n2 + Q * decltype(n1)::value]; });
        Dft<P, INV>::run(t);
        sfor<0, P>([&](auto k1c) {
          constexpr int k1 = decltype(k1c)::value;
          y[k1][n2] = twc<N, k1 * n2, INV>(t[k1]);
        });
      });
      sfor<0, P>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        Dft<Q, INV>::run(y[k1]);
        sfor<0, Q>([&](auto k2) { x[k1 + P * decltype(k2)::value] = y[k1][decltype(k2)::value]; });
      });
    }
  }
};

// -------------------------------------------------------- wave-level sync
// All lanes of one transform live in one wavefront; LDS traffic of one wave
// is processed in order, so a compiler fence + wave barrier orders it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------- cross-lane FFT plan
// FftPlan<T, E, R...>: N = T*E; R... are the cross-lane radices (each divides
// E, product == T).
template <int T_, int E_, int... R>
struct FftPlan {
  static constexpr int T = T_;
  static constexpr int E = E_;
  static constexpr int N = T_ * E_;
  static constexpr int LDS = N + N / E_;  // padded element count
  static_assert(((E_ % R == 0) && ...), "every cross-lane radix must divide E");
  static_assert((R * ... * 1) == T_, "cross-lane radices must multiply to T");
  __device__ __forceinline__ static int pad(int i) { return i + i / E_; }
};

template <class P, int R, int L, bool LAST, bool INV>
__device__ __forceinline__ void cross_stage(float2 (&v)[P::E], float2* buf, const float2* tw, int lane) {
  constexpr int T = P::T, E = P::E, N = P::N;
  constexpr int NB = E / R;          // butterflies per lane
  constexpr int TS = N / (L * R);    // twiddle table stride
  // gather natural-order inputs
  sfor<0, E>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    v[j] = buf[P::pad(lane + T * j)];
  });
  wave_sync();
  sfor<0, NB>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int b = lane + T * q;
    const int jm = b % L;
    float2 u[R];
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      float2 a = v[q + m * NB];
      if constexpr (m > 0) {
        float2 w = tw[jm * m * TS];
        a = INV ? cmulc(a, w) : cmul(a, w);
      }
      u[m] = a;
    });
    Dft<R, INV>::run(u);
    if constexpr (LAST) {
      // out index b + m*N/R == lane + T*(q + m*NB): stays in this lane
      sfor<0, R>([&](auto mc) { v[q + decltype(mc)::value * NB] = u[decltype(mc)::value]; });
    } else {
      const int base = (b - jm) * R + jm;
      sfor<0, R>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        buf[P::pad(base + m * L)] = u[m];
      });
    }
  });
  if constexpr (!LAST) wave_sync();
}

template <class P, bool INV, int L, int R0, int... Rs>
__device__ __forceinline__ void cross_stages(float2 (&v)[P::E], float2* buf, const float2* tw, int lane) {
  cross_stage<P, R0, L, sizeof...(Rs) == 0, INV>(v, buf, tw, lane);
  if constexpr (sizeof...(Rs) > 0) cross_stages<P, INV, L * R0, Rs...>(v, buf, tw, lane);
}

template <class P>
struct FftRun;

template <int T, int E, int... R>
struct FftRun<FftPlan<T, E, R...>> {
  using P = FftPlan<T, E, R...>;
  // In:  v[j] = x[lane + T*j].  Out: v[j] = X[lane + T*j].  Unnormalised.
  // buf: this transform's LDS region (P::LDS entries); tw: W_N^k, k < N.
  template <bool INV>
  __host__ __device__ __forceinline__ static void run(float2 (&v)[E], float2* buf, const float2* tw, int lane) {
    Dft<E, INV>::run(v);
    if constexpr (sizeof...(R) == 0) {
      return;
    } else {
      sfor<0, E>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        buf[P::pad(lane * E + m)] = v[m];
      });
      wave_sync();
      cross_stages<P, INV, E, R...>(v, buf, tw, lane);
    }
  }
};

}  // namespace glsu

// fft_device.h -- wave-level complex FFTs for gfx950 (CDNA4), FP32.
//
// A transform of length N = T * E is owned by T lanes of one wavefront.
// Four-step decomposition (n = n1 + T n2, k = k1 + E k2):
//   X[k1 + E k2] = sum_n1 W_T^{n1 k2} W_N^{n1 k1} sum_n2 x[n1 + T n2] W_E^{n2 k1}
// Stage A: lane n1 holds x[n1 + T n2] in register n2 ("spatial layout") and
//   does the E-point DFT in registers (compile-time twiddles), then multiplies
//   by W_N^{n1 k1} (per-lane constants kept in registers).
// Stage B: E independent T-point DFTs across lanes (batch index k1), done as
//   Stockham radix-R stages (R | E, product of the R's = T) that exchange
//   through LDS laid out as [k1][position].  In every stage a lane owns ONE
//   butterfly position c and E/R values of k1, so its twiddles are a few
//   per-lane constants and every LDS address is lane_base + compile-time
//   offset (no per-element index arithmetic to keep live).  With 64 lanes and
//   a first radix of 4, the first stage's inputs are gathered across the
//   wavefront's four 16-lane rows with permlane swaps instead (FftPlan::XL0).
// The last stage leaves the spectrum in registers ("spectral layout": lane
//   c + (T/R) g, register i + m (E/R) holds X[(g + R i) + E (c + m T/R)]).
// The inverse runs the stages in mirror order (inverse butterflies, conjugate
//   twiddles) from the spectral layout back to the spatial layout, so
//   FFT -> pointwise -> IFFT never reorders data.
// Stockham step (radix r, L = product of earlier radices, butterfly c):
//   in = z[c + m T/r] * W_{L r}^{(c mod L) m};  out = z'[(c div L) L r + c mod L + m L].
// Published algorithms: Cooley-Tukey four-step (Bailey 1990) and the
// Stockham autosort formulation.
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
__host__ __device__ __forceinline__ float2 rmul(float2 a, float h) { return make_float2(a.x * h, a.y * h); }
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

constexpr int ct_gcd(int a, int b) { return b == 0 ? a : ct_gcd(b, a % b); }
constexpr int ct_modinv(int a, int m) {  // a^-1 mod m for gcd(a, m) = 1 (m small)
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return m == 1 ? 0 : -1;
}

// generic: prime sizes by direct O(N^2) sums; composite N = P*Q with gcd(P, Q)
// = 1 by the prime-factor (Good-Thomas) mapping, which needs no twiddles:
//   n = (Q n1 + P n2) mod N,  k = (Q a k1 + P b k2) mod N,  a = Q^-1 mod P, b = P^-1 mod Q,
//   X[k] = sum_n2 W_Q^{n2 k2} sum_n1 x[n] W_P^{n1 k1};
// other composites by one Cooley-Tukey decimation step
//   X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_N^{n2 k1} sum_n1 x[n2 + Q n1] W_P^{n1 k1}.
// Published algorithms: Good (1958) / Thomas (1963), Cooley-Tukey (1965).
template <int N, bool INV>
struct Dft {
  __host__ __device__ __forceinline__ static void run(float2* x) {
    constexpr int P = pick_factor(N);
    constexpr int Q = N / P;
    if constexpr (Q > 1 && ct_gcd(P, Q) == 1) {
      constexpr int a = ct_modinv(Q % P, P), b = ct_modinv(P % Q, Q);
      static_assert(a > 0 && b > 0, "prime-factor mapping needs coprime factors");
      float2 y[P][Q];
      sfor<0, Q>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        float2 t[P];
        sfor<0, P>([&](auto n1) { t[decltype(n1)::value] = x[(Q * decltype(n1)::value + P * n2) % N]; });
        Dft<P, INV>::run(t);
        sfor<0, P>([&](auto k1c) { y[decltype(k1c)::value][n2] = t[decltype(k1c)::value]; });
      });
      sfor<0, P>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        Dft<Q, INV>::run(y[k1]);
        sfor<0, Q>([&](auto k2c) {
          constexpr int k2 = decltype(k2c)::value;
          x[(Q * a * k1 + P * b * k2) % N] = y[k1][k2];
        });
      });
    } else if constexpr (Q == 1) {
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

// ---------------------------------------------------------------- plans
template <int... R>
struct RadixList {
  static constexpr int n = sizeof...(R);
  static constexpr int v[n] = {R...};
  static constexpr int at(int s) { return v[s]; }
  static constexpr int prod_before(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= v[i];
    return p;
  }
  static constexpr int maxr() {
    int m = 1;
    for (int i = 0; i < n; ++i) m = v[i] > m ? v[i] : m;
    return m;
  }
  static constexpr bool uniform_pow2() {
    for (int i = 0; i < n; ++i)
      if (v[i] != v[0]) return false;
    return (v[0] & (v[0] - 1)) == 0;
  }
};

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// FftPlan<T, E, R...>: N = T*E; R... are the cross-lane radices (each divides
// E, product == T).  LDS per transform: E rows of ROW elements.
//
// Bank-conflict-free layout for uniform power-of-two radices (T = R^NS): a
// position p with base-R digits (d_{NS-1} ... d_0) is stored at
//   R^{NS-1} d_{NS-1} + sum_{j < NS-1} ((d_j + d_{NS-1}) mod R) R^j,
// and rows are ROW = T + T/R apart.  Reads (p = c + m T/R), the scattered
// Stockham writes (p = (c div L) L R + c mod L + m L) and the transition
// accesses then touch distinct banks across every 16- or 32-lane group
// (checked exhaustively for the compiled plans); without the skew the stage
// writes at stride R conflict 4-ways.
template <int T_, int E_, int... R>
struct FftPlan {
  static constexpr int T = T_;
  static constexpr int E = E_;
  static constexpr int N = T_ * E_;
  using Rad = RadixList<R...>;
  static constexpr int NS = sizeof...(R);
  static constexpr int RL = Rad::at(NS - 1);  // last radix
  static constexpr bool SKEW = Rad::uniform_pow2();
  static constexpr int ROW = SKEW ? T_ + T_ / Rad::at(0) : T_ + 4;  // row pitch (elements)
  static constexpr int LDS = E_ * ROW;
  __device__ __forceinline__ static int phys(int p) {
    if constexpr (!SKEW || NS == 1) {
      return p;
    } else {
      constexpr int R0 = Rad::at(0), LB = ilog2(R0), TS = LB * (NS - 1);
      const int top = p >> TS;
      int out = top << TS;
#pragma unroll
      for (int j = 0; j < NS - 1; ++j) out |= (((p >> (LB * j)) + top) & (R0 - 1)) << (LB * j);
      return out;
    }
  }
  static_assert(((E_ % R == 0) && ...), "every cross-lane radix must divide E");
  static_assert((R * ... * 1) == T_, "cross-lane radices must multiply to T");
  // spectral layout: k of register j in lane (c, g) = kbase(lane) + koff(j)
  __host__ __device__ __forceinline__ static int kbase(int lane) {
    const int c = lane % (T_ / RL), g = lane / (T_ / RL);
    return g + E_ * c;
  }
  static constexpr int koff(int j) {
    const int nb = E_ / RL, i = j % nb, m = j / nb;
    return RL * i + E_ * (T_ / RL) * m;
  }
  // The transition into the first cross-lane stage (and, in the inverse, out
  // of it) moves data between the wavefront's four 16-lane rows in registers
  // (v_permlane32_swap / v_permlane16_swap, xl_gather / xl_scatter) instead of
  // an LDS write + read: one LDS round trip fewer per transform, the same
  // arithmetic.  Needs a whole wavefront per transform and a first radix of 4
  // (the stage's lane groups are then exactly the rows).
  static constexpr bool XL0 = T_ == 64 && Rad::at(0) == 4 && NS > 1;
};

// ----------------------------------------------- row exchange in registers
// Swap of lanes 32..63 of a with lanes 0..31 of b (v_permlane32_swap), and of
// the odd 16-lane rows of a with the even rows of b (v_permlane16_swap).
__device__ __forceinline__ void swap_half32(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_rows16(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}
// 4 x 4 transpose between the rows of the wavefront (row g = lanes 16g..16g+15)
// and four registers: afterwards register m of row g holds what register g of
// row m held (lane within the row unchanged).  The transpose is its own inverse.
__device__ __forceinline__ void xpose_rows4(float2& r0, float2& r1, float2& r2, float2& r3) {
  swap_half32(r0.x, r2.x);
  swap_half32(r0.y, r2.y);
  swap_half32(r1.x, r3.x);
  swap_half32(r1.y, r3.y);
  swap_rows16(r0.x, r1.x);
  swap_rows16(r0.y, r1.y);
  swap_rows16(r2.x, r3.x);
  swap_rows16(r2.y, r3.y);
}

// Stage-0 gather in registers (forward): lane n1 = c + 16 g' holds the
// twiddled x[k1] of its column n1 in register k1.  Stage 0 wants, in lane
// (c, g), register i + m NB = element k1 = g + 4 i of column c + 16 m.
// Group i (registers 4i .. 4i+3) transposed across rows gives exactly that.
template <class P>
__device__ __forceinline__ void xl_gather(float2 (&v)[P::E]) {
  constexpr int E = P::E, NB = E / 4;
  float2 w[E];
  sfor<0, NB>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    xpose_rows4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    sfor<0, 4>([&](auto mc) { w[i + decltype(mc)::value * NB] = v[4 * i + decltype(mc)::value]; });
  });
  sfor<0, E>([&](auto kc) { v[decltype(kc)::value] = w[decltype(kc)::value]; });
}
// Inverse of xl_gather.
template <class P>
__device__ __forceinline__ void xl_scatter(float2 (&v)[P::E]) {
  constexpr int E = P::E, NB = E / 4;
  float2 w[E];
  sfor<0, NB>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    float2 r[4];
    sfor<0, 4>([&](auto mc) { r[decltype(mc)::value] = v[i + decltype(mc)::value * NB]; });
    xpose_rows4(r[0], r[1], r[2], r[3]);
    sfor<0, 4>([&](auto mc) { w[4 * i + decltype(mc)::value] = r[decltype(mc)::value]; });
  });
  sfor<0, E>([&](auto kc) { v[decltype(kc)::value] = w[decltype(kc)::value]; });
}

// Per-lane constants of one plan: transition twiddles W_N^{lane k1} and the
// stage twiddles W_{L R}^{(c mod L) m}, read once from the W_N^k table.
// TT_LDS: the transition twiddles live in a per-workgroup LDS table
// ([k1][lane], shared by every transform of the workgroup) instead of E VGPR
// pairs; kernels that hold several spectra in registers use it.
template <class P, bool TT_LDS = false>
struct FftLane {
  static constexpr int T = P::T, E = P::E, N = P::N, NS = P::NS;
  static constexpr int RM = P::Rad::maxr();
  static constexpr bool kTTLds = TT_LDS;
  int lane;
  float2 tt[TT_LDS ? 1 : E];
  const float2* ttab;  // LDS [E][T] table when TT_LDS
  float2 ts[NS][RM];
  // TT_LDS: ttab_lds must hold E*T entries; the whole workgroup must call
  // fill_ttab + __syncthreads before any transform.
  __device__ __forceinline__ static void fill_ttab(float2* ttab_lds, const float2* __restrict__ twN) {
    for (int i = threadIdx.x; i < E * T; i += blockDim.x) {
      const int k = i / T, l = i % T;
      ttab_lds[i] = twN[(l * k) % N];
    }
  }
  // Transition twiddles of one transform: tw_row() once per transform, then
  // tw_trans(row, k) for each k.  With TT_LDS the row pointer goes through an
  // opaque copy of the lane index, so every transform re-reads the table (the
  // compiler cannot keep E twiddle pairs live across transforms) while the E
  // reads of one transform stay independent and issue back to back.
  __device__ __forceinline__ const float2* tw_row() const {
    if constexpr (TT_LDS) {
      int o = lane;
      __asm__ volatile("" : "+v"(o));
      return ttab + o;
    } else {
      return nullptr;
    }
  }
  __device__ __forceinline__ float2 tw_trans(const float2* row, int k) const {
    if constexpr (TT_LDS) {
      return row[k * T];
    } else {
      return tt[k];
    }
  }
  __device__ __forceinline__ void init(int lane_, const float2* __restrict__ twN,
                                       const float2* ttab_lds = nullptr) {
    lane = lane_;
    ttab = ttab_lds;
    if constexpr (!TT_LDS) {
#pragma unroll
      for (int k = 0; k < E; ++k) tt[k] = twN[(lane * k) % N];
    }
    sfor<0, NS>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      constexpr int R = P::Rad::at(s), L = P::Rad::prod_before(s);
      const int c = lane % (T / R);
#pragma unroll
      for (int m = 0; m < R; ++m) ts[s][m] = twN[((c % L) * m * (N / (L * R))) % N];
    });
  }
};

template <class P, int S, bool INV, class Ctx, bool REG_IN = false>
__device__ __forceinline__ void stage_fwd(float2 (&v)[P::E], float2* buf, const Ctx& L_) {
  // forward Stockham stage S: read z[k1][c + m T/R] -> twiddle -> DFT_R -> write
  // (REG_IN: the stage's inputs are already in v, gathered in registers)
  constexpr int T = P::T, E = P::E, ROW = P::ROW;
  constexpr int R = P::Rad::at(S), L = P::Rad::prod_before(S), NB = E / R;
  constexpr bool LAST = (S == P::NS - 1);
  const int lane = L_.lane;
  const int c = lane % (T / R), g = lane / (T / R);
  if constexpr (!REG_IN) {
    const float2* rb = buf + g * ROW;
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const int o = P::phys(c + m * (T / R));
      sfor<0, NB>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v[i + m * NB] = rb[R * i * ROW + o];
      });
    });
  }
  wave_sync();
  float2* wb = buf + g * ROW;
  const int wbase = (c / L) * L * R + (c % L);
  sfor<0, NB>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    float2 u[R];
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      float2 a = v[i + m * NB];
      if constexpr (L > 1 && m > 0) a = INV ? cmulc(a, L_.ts[S][m]) : cmul(a, L_.ts[S][m]);
      u[m] = a;
    });
    Dft<R, INV>::run(u);
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if constexpr (LAST) v[i + m * NB] = u[m];
      else wb[R * i * ROW + P::phys(wbase + m * L)] = u[m];
    });
  });
  if constexpr (!LAST) wave_sync();
}

template <class P, int S, class Ctx, bool REG_OUT = false>
__device__ __forceinline__ void stage_inv(float2 (&v)[P::E], float2* buf, const Ctx& L_) {
  // mirror of stage_fwd: read outputs -> inverse DFT_R -> conj twiddle -> write inputs
  // (REG_OUT: the inputs stay in v, in the stage's register order, for xl_scatter)
  constexpr int T = P::T, E = P::E, ROW = P::ROW;
  constexpr int R = P::Rad::at(S), L = P::Rad::prod_before(S), NB = E / R;
  constexpr bool LAST = (S == P::NS - 1);
  const int lane = L_.lane;
  const int c = lane % (T / R), g = lane / (T / R);
  if constexpr (!LAST) {
    const float2* rb = buf + g * ROW;
    const int rbase = (c / L) * L * R + (c % L);
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const int o = P::phys(rbase + m * L);
      sfor<0, NB>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v[i + m * NB] = rb[R * i * ROW + o];
      });
    });
    wave_sync();
  }
  float2* wb = buf + g * ROW;
  sfor<0, NB>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    float2 u[R];
    sfor<0, R>([&](auto mc) { u[decltype(mc)::value] = v[i + decltype(mc)::value * NB]; });
    Dft<R, true>::run(u);
    sfor<0, R>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      float2 a = u[m];
      if constexpr (L > 1 && m > 0) a = cmulc(a, L_.ts[S][m]);
      if constexpr (REG_OUT) v[i + m * NB] = a;
      else wb[R * i * ROW + P::phys(c + m * (T / R))] = a;
    });
  });
  if constexpr (!REG_OUT) wave_sync();
}

template <class P, int S, class Ctx, int LO = 0>
__device__ __forceinline__ void stages_inv(float2 (&v)[P::E], float2* buf, const Ctx& L_) {
  stage_inv<P, S>(v, buf, L_);
  if constexpr (S > LO) stages_inv<P, S - 1, Ctx, LO>(v, buf, L_);
}

// Forward DFT (unnormalised): spatial layout in, spectral layout out.
template <class P, class Ctx>
__device__ __forceinline__ void fft_fwd(float2 (&v)[P::E], float2* buf, const Ctx& L_) {
  constexpr int E = P::E, ROW = P::ROW;
  // LDS twiddles: issue the table reads before the in-register DFT, whose
  // arithmetic then covers their latency (the scheduling barrier keeps the
  // scheduler from sinking them back to their uses)
  float2 tw[E];
  const float2* tr = L_.tw_row();
  if constexpr (Ctx::kTTLds) {
    sfor<1, E>([&](auto kc) { tw[decltype(kc)::value] = L_.tw_trans(tr, decltype(kc)::value); });
    __builtin_amdgcn_sched_barrier(0);
  }
  Dft<E, false>::run(v);
  if constexpr (P::XL0) {
    sfor<1, E>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      v[k] = cmul(v[k], Ctx::kTTLds ? tw[k] : L_.tw_trans(tr, k));
    });
    xl_gather<P>(v);
    stage_fwd<P, 0, false, Ctx, true>(v, buf, L_);
    sfor<1, P::NS>([&](auto sc) { stage_fwd<P, decltype(sc)::value, false>(v, buf, L_); });
  } else {
    float2* wb = buf + P::phys(L_.lane);
    sfor<0, E>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (k == 0) {
        wb[0] = v[0];
      } else {
        wb[k * ROW] = cmul(v[k], Ctx::kTTLds ? tw[k] : L_.tw_trans(tr, k));
      }
    });
    wave_sync();
    sfor<0, P::NS>([&](auto sc) { stage_fwd<P, decltype(sc)::value, false>(v, buf, L_); });
  }
}

// Inverse DFT (unnormalised, e^{+i}): spectral layout in, spatial layout out.
template <class P, class Ctx>
__device__ __forceinline__ void fft_inv(float2 (&v)[P::E], float2* buf, const Ctx& L_) {
  constexpr int E = P::E, ROW = P::ROW;
  if constexpr (P::NS > 1) stages_inv<P, P::NS - 1, Ctx, 1>(v, buf, L_);
  // LDS twiddles: issued before the last stage, whose exchange and
  // butterflies cover their latency
  float2 tw[E];
  const float2* tr = L_.tw_row();
  if constexpr (Ctx::kTTLds) {
    sfor<1, E>([&](auto kc) { tw[decltype(kc)::value] = L_.tw_trans(tr, decltype(kc)::value); });
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (P::XL0) {
    stage_inv<P, 0, Ctx, true>(v, buf, L_);
    xl_scatter<P>(v);
    sfor<1, E>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      v[k] = cmulc(v[k], Ctx::kTTLds ? tw[k] : L_.tw_trans(tr, k));
    });
  } else {
    stage_inv<P, 0>(v, buf, L_);
    const float2* rb = buf + P::phys(L_.lane);
    sfor<0, E>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      const float2 a = rb[k * ROW];
      if constexpr (k == 0) {
        v[0] = a;
      } else {
        v[k] = cmulc(a, Ctx::kTTLds ? tw[k] : L_.tw_trans(tr, k));
      }
    });
  }
  Dft<E, true>::run(v);
}

}  // namespace glsu

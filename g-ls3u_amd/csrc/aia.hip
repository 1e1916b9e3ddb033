// aia.hip -- AIA (advanced iterative algorithm; Z. Wang and B. Han, Opt. Lett.
// 29, 1671 (2004)) on gfx950: the initial phase Phi0 and the phase shifts of
// M >= 3 randomly phase-shifted frames.  Replaces AIA::AIA_CUDAF
// (AIA/aia_cudaf.h:15-72, AIA/aia_cudaf.cu:204-461) and the CPU classes of the
// same algorithm (AIA/aia_cpuf.cpp, AIA/aia_cpu.cpp; MatLab/fp_aia.m:33-121).
//
// Model: I_k(j) = a_j + b_j cos(phi_j + delta_k).  Each iteration alternates
//   phase step  x_j = A1^-1 sum_k I_k(j) v_k,   v_k = (1, cos d_k, sin d_k),
//               phi_j = atan2(-x2, x1);
//   shift step  y_k = A2^-1 sum_j I_k(j) u_j,   u_j = (1, cos phi_j, sin phi_j),
//               d_k = atan2(-y2, y1).
// A1 is one matrix for all pixels, so its solve folds into per-frame weights
// w_k = A1^-1 v_k: x1 = sum_k I_k w1_k, x2 = sum_k I_k w2_k (the reference
// assembles a block-diagonal sparse system of 3N unknowns and factors it with
// cuSOLVER every iteration, aia_cudaf.cu:417-426).  The weights of each row sum
// to zero (sum_k w_k = A1^-1 A1 e0 = e0), so the pixel's first frame is
// subtracted before the sum, which removes the background from the FP32
// accumulation.  cos phi = x1/r and sin phi = -x2/r with r = |(x1, x2)|.
//
// An iteration is two launches.  aia_pass_kernel does the phase step and
// accumulates the sums of the shift step (sum_j u_j u_j^T and sum_j I_k(j)
// (cos phi_j, sin phi_j)) from the same frame reads (the reference reads every
// frame M+1 times per iteration, aia_cudaf.cu:424 and 433-438); its blocks write
// FP64 partials.  aia_solve_kernel (one block) reduces them in a fixed order,
// solves the M shift systems and the next A1 in FP64, tests convergence and
// writes the next weights.  (A last-block-done reduction inside the pass would
// need an agent-scope release/acquire per block, which on this multi-XCD part
// writes back and invalidates the XCD's L2 -- evicting the frames the other
// blocks are reading.)  A device flag makes the remaining launches exit at
// once after convergence, so iterations need no host round trip; results are
// bitwise reproducible run to run.
//
// Frames with more than 32 images are processed in chunks of 32 frames per
// block group (each group recomputes the phase of its pixels and accumulates
// the shift sums of its own frames).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/glsu.h"

namespace glsu {
void set_error(const char* msg);
}

namespace {

constexpr int kThreads = 256;
constexpr int kSolveThreads = 1024;  // the one-block shift-step solve
constexpr int kMaxFrames = 1024;
constexpr int kLaunchGroup = 16;  // iterations enqueued between convergence polls

struct AiaCtl {
  int done;      // 1 once converged, out of iterations, or failed
  int iters;     // completed iterations
  int status;    // 0, or 10*step + k: leading minor k not positive definite
                 // (step 1: phase step matrix A1, step 2: shift step matrix A2)
  int pad;
  double err;
};

int afail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  glsu::set_error(buf);
  return code;
}

#define ATRY(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return afail(GLSU_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---- device helpers ---------------------------------------------------------
// Wave sum with DPP lane moves (no LDS traffic): butterflies within quads,
// half-row and row mirrors give every lane its row-of-16 sum; row_bcast:15 and
// row_bcast:31 then fold rows 0-2 into row 3, and lane 63 holds the total.
// Every lane of the wave must be active.
#define GLSU_DPP(v, ctrl, rows)                                                              \
  __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, \
                                                        rows, 0xF, false))
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += GLSU_DPP(v, 0xB1, 0xF);   // quad_perm [1,0,3,2]
  v += GLSU_DPP(v, 0x4E, 0xF);   // quad_perm [2,3,0,1]
  v += GLSU_DPP(v, 0x141, 0xF);  // row_half_mirror
  v += GLSU_DPP(v, 0x140, 0xF);  // row_mirror
  v += GLSU_DPP(v, 0x142, 0xA);  // row_bcast:15 into rows 1 and 3
  v += GLSU_DPP(v, 0x143, 0xC);  // row_bcast:31 into rows 2 and 3
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
#undef GLSU_DPP
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Four consecutive pixels of one frame as floats; pixels at or beyond N read 0.
template <typename T>
__device__ __forceinline__ void load4(const T* __restrict__ fk, long long p0, long long N, bool vec,
                                      float v[4]);
template <>
__device__ __forceinline__ void load4<uint8_t>(const uint8_t* __restrict__ fk, long long p0,
                                               long long N, bool vec, float v[4]) {
  if (vec) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(fk + p0);
    v[0] = (float)(w & 255u);
    v[1] = (float)((w >> 8) & 255u);
    v[2] = (float)((w >> 16) & 255u);
    v[3] = (float)(w >> 24);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = p0 + i < N ? (float)fk[p0 + i] : 0.f;
  }
}
template <>
__device__ __forceinline__ void load4<float>(const float* __restrict__ fk, long long p0,
                                             long long N, bool vec, float v[4]) {
  if (vec) {
    const float4 w = *reinterpret_cast<const float4*>(fk + p0);
    v[0] = w.x; v[1] = w.y; v[2] = w.z; v[3] = w.w;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = p0 + i < N ? fk[p0 + i] : 0.f;
  }
}
template <>
__device__ __forceinline__ void load4<double>(const double* __restrict__ fk, long long p0,
                                              long long N, bool vec, float v[4]) {
  if (vec) {
    const double2 a = *reinterpret_cast<const double2*>(fk + p0);
    const double2 b = *reinterpret_cast<const double2*>(fk + p0 + 2);
    v[0] = (float)a.x; v[1] = (float)a.y; v[2] = (float)b.x; v[3] = (float)b.y;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = p0 + i < N ? (float)fk[p0 + i] : 0.f;
  }
}

// Cholesky A = U^T U of the symmetric 3x3 a = (a00, a01, a02, a11, a12, a22),
// as LAPACKE_?posv 'U' factors it (AIA/aia_cpuf.cpp:161, 238).  Returns 0, or
// the order of the first leading minor that is not positive definite.  A pivot
// at or below 1e-10 of its diagonal entry counts as not positive: rounding can
// leave a singular matrix (all shifts equal) a tiny positive pivot.
constexpr double kCholTol = 1e-10;
__host__ __device__ int chol3(const double a[6], double u[6]) {
  if (!(a[0] > 0)) return 1;
  u[0] = sqrt(a[0]);
  u[1] = a[1] / u[0];
  u[2] = a[2] / u[0];
  const double d1 = a[3] - u[1] * u[1];
  if (!(d1 > kCholTol * a[3])) return 2;
  u[3] = sqrt(d1);
  u[4] = (a[4] - u[1] * u[2]) / u[3];
  const double d2 = a[5] - u[2] * u[2] - u[4] * u[4];
  if (!(d2 > kCholTol * a[5])) return 3;
  u[5] = sqrt(d2);
  return 0;
}
__host__ __device__ void chol3_solve(const double u[6], double b[3]) {
  const double y0 = b[0] / u[0];
  const double y1 = (b[1] - u[1] * y0) / u[3];
  const double y2 = (b[2] - u[2] * y0 - u[4] * y1) / u[5];
  const double x2 = y2 / u[5];
  const double x1 = (y1 - u[4] * x2) / u[3];
  b[0] = (y0 - u[1] * x1 - u[2] * x2) / u[0];
  b[1] = x1;
  b[2] = x2;
}

// Phase-step weights from the shifts: A1 = sum_k v_k v_k^T in FP64,
// w_k = A1^-1 v_k, rounded to float (the solve kernel computes the same on the
// device after every iteration).  Returns the Cholesky status.
int phase_weights_host(const double* delta, int M, float* w1, float* w2) {
  double t[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < M; ++k) {
    const double c = std::cos(delta[k]), s = std::sin(delta[k]);
    t[0] += c; t[1] += c * c; t[2] += s; t[3] += c * s; t[4] += s * s;
  }
  const double a[6] = {(double)M, t[0], t[2], t[1], t[3], t[4]};
  double u[6];
  const int info = chol3(a, u);
  if (info) return info;
  for (int k = 0; k < M; ++k) {
    double b[3] = {1.0, std::cos(delta[k]), std::sin(delta[k])};
    chol3_solve(u, b);
    w1[k] = (float)b[1];
    w2[k] = (float)b[2];
  }
  return 0;
}

// Frames padded to a multiple of 32 in the shared weight arrays (zeros past M).
__host__ __device__ __forceinline__ int padded_frames(int M) { return (M + 31) & ~31; }

// Four uint8 pixels of one frame as one packed word (byte i = pixel p0 + i);
// pixels at or beyond N read 0.
__device__ __forceinline__ uint32_t load_raw4(const uint8_t* __restrict__ fk, long long p0,
                                              long long N, bool vec) {
  if (vec) return *reinterpret_cast<const uint32_t*>(fk + p0);
  uint32_t w = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (p0 + i < N) w |= (uint32_t)fk[p0 + i] << (8 * i);
  return w;
}

// Phase-step sums over frames kb .. kb+GS-1 of one quad.  All GS loads are
// unconditional (indices clamped to M-1; the padded weights are zero) so they
// are in flight together; uint8 frames are loaded as packed words (one
// register per frame).  keep (uint8 only, may be null): the packed words.
template <typename T, int GS>
__device__ __forceinline__ void phase_group(const T* __restrict__ f, int kb, int M, long long N,
                                            long long p0, bool vec, const float* sw1,
                                            const float* sw2, const float i0[4], float x1[4],
                                            float x2[4], uint32_t* keep) {
  if constexpr (sizeof(T) == 1) {
    uint32_t r[GS];
#pragma unroll
    for (int j = 0; j < GS; ++j) {
      const int kk = kb + j < M ? kb + j : M - 1;
      r[j] = load_raw4(reinterpret_cast<const uint8_t*>(f) + (size_t)kk * N, p0, N, vec);
    }
#pragma unroll
    for (int j = 0; j < GS; ++j) {
      const float a = sw1[kb + j], b = sw2[kb + j];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const float d = (float)((r[j] >> (8 * p)) & 255u) - i0[p];
        x1[p] = fmaf(d, a, x1[p]);
        x2[p] = fmaf(d, b, x2[p]);
      }
      if (keep) keep[j] = r[j];
    }
  } else {
    float v[GS][4];
#pragma unroll
    for (int j = 0; j < GS; ++j) {
      const int kk = kb + j < M ? kb + j : M - 1;
      load4<T>(f + (size_t)kk * N, p0, N, vec, v[j]);
    }
#pragma unroll
    for (int j = 0; j < GS; ++j) {
      const float a = sw1[kb + j], b = sw2[kb + j];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const float d = v[j][p] - i0[p];
        x1[p] = fmaf(d, a, x1[p]);
        x2[p] = fmaf(d, b, x2[p]);
      }
    }
  }
}

// ---- kernels ------------------------------------------------------------------
// Per-frame intensity sums sum_j I_k(j) (constant over the iterations): block
// (b, k) sums quads b*256 + tid (stride Gs*256) of frame k in FP64.  Block
// (0, 0) also initialises the control block (the shifts and the first
// weights come from the host).
template <typename T>
__global__ __launch_bounds__(kThreads) void aia_frame_sums_kernel(
    const T* __restrict__ f, long long N, int vec, double* __restrict__ part,
    AiaCtl* __restrict__ ctl, int max_iter, double max_err, int timing) {
  const int k = blockIdx.y, Gs = gridDim.x;
  if (blockIdx.x == 0 && k == 0 && threadIdx.x == 0) {
    const double err = max_err * 2.0;
    ctl->iters = 0;
    ctl->err = err;
    ctl->status = 0;
    ctl->done = timing ? max_iter <= 0 : !(err > max_err && 0 < max_iter);
  }
  const T* fk = f + (size_t)k * N;
  const long long nq = (N + 3) >> 2;
  const long long S = (long long)Gs * kThreads;
  double s = 0;
  for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < nq; q += 8 * S) {
    float v[8][4];
#pragma unroll
    for (int u = 0; u < 8; ++u) {  // eight quads in flight; out-of-range ones re-read q
      const long long qq = q + u * S < nq ? q + u * S : q;
      load4<T>(fk, qq << 2, N, vec, v[u]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q + u * S < nq) s += ((double)v[u][0] + (double)v[u][1]) + ((double)v[u][2] + (double)v[u][3]);
  }
  __shared__ double red[kThreads / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    part[(size_t)k * Gs + blockIdx.x] = t;
  }
}

// Phase step + shift-step sums of one iteration.  Grid: G * nchunks blocks;
// block (c, b) = blockIdx.x / G, % G handles the pixel quads q = b*256 + tid
// (stride G*256) and the shift sums of frames c*KC .. c*KC+KC-1.  Partials:
// part[v * Gtot + block], v = 0..4 the five sums of (cos, sin) products
// (chunk 0 only), 5 + i / 5 + KC + i the sums of I_k cos phi / I_k sin phi of
// frame c*KC + i, each the FP64 sum of the block's four waves rounded to
// FP32 (FP64 again across blocks).  No cross-block synchronisation: aia_solve_kernel, the next
// launch on the stream, reduces the partials.
template <typename T, int KC>
__global__ __launch_bounds__(kThreads) void aia_pass_kernel(
    const T* __restrict__ f, int M, long long N, int vec, int G, const AiaCtl* __restrict__ ctl,
    const float* __restrict__ w1g, const float* __restrict__ w2g, float* __restrict__ part) {
  extern __shared__ unsigned char smem_raw[];
  float* sw1 = reinterpret_cast<float*>(smem_raw);  // [padded_frames(M)], zero past M
  float* sw2 = sw1 + padded_frames(M);              // [padded_frames(M)]
  __shared__ double s_red[(kThreads / 64) * (5 + 2 * KC)];

  if (ctl->done) return;  // uniform: written by the previous solve launch
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int c = blockIdx.x / G, b = blockIdx.x - c * G;
  const int Gtot = gridDim.x;
  const int k0 = c * KC;
  const int Mp = padded_frames(M);
  for (int k = tid; k < Mp; k += kThreads) {
    sw1[k] = k < M ? w1g[k] : 0.f;
    sw2[k] = k < M ? w2g[k] : 0.f;
  }
  __syncthreads();

  constexpr bool kKeep = sizeof(T) == 1;  // u8: keep this chunk's frames in registers
  float acc[5] = {0, 0, 0, 0, 0};
  float P[KC], Q[KC];
#pragma unroll
  for (int i = 0; i < KC; ++i) P[i] = Q[i] = 0.f;

  const long long nq = (N + 3) >> 2;
  for (long long q = (long long)b * kThreads + tid; q < nq; q += (long long)G * kThreads) {
    const long long p0 = q << 2;
    float i0[4], x1[4] = {0, 0, 0, 0}, x2[4] = {0, 0, 0, 0};
    uint32_t keep[kKeep ? KC : 1];
    load4<T>(f, p0, N, vec, i0);
    // this chunk's frames, then the others, eight at a time (groups of eight
    // never straddle a chunk boundary: KC is a multiple of 8)
    if constexpr (kKeep) {
      phase_group<T, KC>(f, k0, M, N, p0, vec, sw1, sw2, i0, x1, x2, keep);
    } else {
#pragma unroll
      for (int g = 0; g < KC / 8; ++g)
        if (k0 + 8 * g < M)
          phase_group<T, 8>(f, k0 + 8 * g, M, N, p0, vec, sw1, sw2, i0, x1, x2, nullptr);
    }
    for (int kb = 0; kb < M; kb += 8) {
      if (kb >= k0 && kb < k0 + KC) continue;
      phase_group<T, 8>(f, kb, M, N, p0, vec, sw1, sw2, i0, x1, x2, nullptr);
    }
    float cp[4], sp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float r2 = x1[p] * x1[p] + x2[p] * x2[p];
      if (p0 + p >= N) {
        cp[p] = sp[p] = 0.f;  // outside the image: contributes nothing
      } else if (r2 > 0.f) {
        const float inv = rsqrtf(r2);
        cp[p] = x1[p] * inv;
        sp[p] = -x2[p] * inv;
      } else {
        cp[p] = 1.f;  // phi = atan2(0, 0) = 0
        sp[p] = 0.f;
      }
    }
    if (c == 0) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        acc[0] += cp[p];
        acc[1] = fmaf(cp[p], cp[p], acc[1]);
        acc[2] += sp[p];
        acc[3] = fmaf(cp[p], sp[p], acc[3]);
        acc[4] = fmaf(sp[p], sp[p], acc[4]);
      }
    }
#pragma unroll
    for (int g = 0; g < KC / 8; ++g) {
      if (k0 + 8 * g < M) {
        float v[8][4];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if constexpr (kKeep) {
            // re-unpack from the packed word (the barrier stops the compiler
            // from keeping the phase step's 4*KC unpacked floats alive)
            uint32_t w = keep[8 * g + j];
            asm volatile("" : "+v"(w));
            v[j][0] = (float)(w & 255u);
            v[j][1] = (float)((w >> 8) & 255u);
            v[j][2] = (float)((w >> 16) & 255u);
            v[j][3] = (float)(w >> 24);
          } else {
            const int kk = k0 + 8 * g + j < M ? k0 + 8 * g + j : M - 1;
            load4<T>(f + (size_t)kk * N, p0, N, vec, v[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float pc = 0.f, ps = 0.f;
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            pc = fmaf(v[j][p], cp[p], pc);
            ps = fmaf(v[j][p], sp[p], ps);
          }
          P[8 * g + j] += pc;  // slots past M are never written out
          Q[8 * g + j] += ps;
        }
      }
    }
  }

  // block partials (FP32 within a wave, FP64 across waves and blocks); frame
  // slots past M are never read and are skipped (a uniform branch)
  constexpr int V = 5 + 2 * KC;
  if (c == 0) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const float s = wave_sum_dpp(acc[i]);
      if (lane == 0) s_red[wid * V + i] = s;
    }
  }
#pragma unroll
  for (int i = 0; i < KC; ++i) {
    if (k0 + i < M) {
      const float sp_ = wave_sum_dpp(P[i]);
      const float sq_ = wave_sum_dpp(Q[i]);
      if (lane == 0) {
        s_red[wid * V + 5 + i] = sp_;
        s_red[wid * V + 5 + KC + i] = sq_;
      }
    }
  }
  __syncthreads();
  const int nk = M - k0 < KC ? M - k0 : KC;
  const bool live = tid < 5 ? c == 0 : (tid < 5 + KC ? tid - 5 < nk : (tid < V && tid - 5 - KC < nk));
  if (live) {
    double t = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += s_red[w * V + tid];
    part[(size_t)tid * Gtot + blockIdx.x] = (float)t;
  }
}

// Shift step of one iteration (one block): reduces the pass's partials in a
// fixed order, solves the M shift systems in FP64, tests convergence, and
// writes the next phase-step weights and the control block.  After the done
// flag, every global load is issued at once (shifts, frame sums, partials) so
// the block waits on memory once more; the rest works from LDS.
template <int KC>
__global__ __launch_bounds__(kSolveThreads) void aia_solve_kernel(
    int M, long long N, int G, int Gtot, AiaCtl* __restrict__ ctl, float* __restrict__ w1g,
    float* __restrict__ w2g, double* __restrict__ delta, double* __restrict__ Tk,
    const float* __restrict__ part, const double* __restrict__ sums_part, int Gs_first,
    int max_iter, double max_err, int err_mode, int timing) {
  constexpr int kWaves = kSolveThreads / 64;
  extern __shared__ unsigned char smem_raw[];
  double* tot = reinterpret_cast<double*>(smem_raw);  // [5 + 2M]: five sums, then P_k, Q_k
  double* dold = tot + 5 + 2 * M;                      // [M]
  double* dnew = dold + M;                             // [M]
  double* tk = dnew + M;                               // [M]
  double* cs = tk + M;                                 // [2M] cos, sin of the new shifts
  __shared__ double s_red[5 * kWaves];
  __shared__ double s_err[kWaves];
  __shared__ double s_u[6];
  __shared__ int s_info, s_iters;
  if (ctl->done) return;  // converged earlier: one load, nothing written
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_iters = ctl->iters;
  for (int k = tid; k < M; k += kSolveThreads) {
    dold[k] = delta[k];
    if (Gs_first == 0) tk[k] = Tk[k];
  }
  if (Gs_first > 0)  // first iteration: the frame sums, one wave per frame
    for (int k = wid; k < M; k += kWaves) {
      const double t = wave_sum(lane < Gs_first ? sums_part[(size_t)k * Gs_first + lane] : 0.0);
      if (lane == 0) tk[k] = t;
    }
  const int nv = 5 + 2 * M;
  // each wave reduces four values at a time; each lane keeps up to 64 loads in
  // flight (four values x sixteen block strides: one round for up to 1024
  // pass blocks), clamped in range and masked
  for (int v0 = wid * 4; v0 < nv; v0 += 4 * kWaves) {
    size_t base[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = v0 + j < nv ? v0 + j : nv - 1;
      int row, g0;
      if (v < 5) {
        row = v;
        g0 = 0;
      } else {
        const int k = (v - 5) >> 1, cc = k / KC, i = k - cc * KC;
        row = 5 + i + (((v - 5) & 1) ? KC : 0);
        g0 = cc * G;
      }
      base[j] = (size_t)row * Gtot + g0;
    }
    double s[4] = {0, 0, 0, 0};
    for (int g = lane; g < G; g += 1024) {
      float t[16][4];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int gg = g + 64 * u < G ? g + 64 * u : g;
#pragma unroll
        for (int j = 0; j < 4; ++j) t[u][j] = part[base[j] + gg];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (g + 64 * u < G)
#pragma unroll
          for (int j = 0; j < 4; ++j) s[j] += (double)t[u][j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double r = wave_sum(s[j]);
      if (lane == 0 && v0 + j < nv) tot[v0 + j] = r;
    }
  }
  __syncthreads();
  if (Gs_first > 0)
    for (int k = tid; k < M; k += kSolveThreads) Tk[k] = tk[k];
  if (tid == 0) {
    const double a[6] = {(double)N, tot[0], tot[2], tot[1], tot[3], tot[4]};
    s_info = chol3(a, s_u);
  }
  __syncthreads();
  if (s_info) {
    if (tid == 0) {
      ctl->status = 20 + s_info;
      ctl->done = 1;
    }
    return;
  }
  for (int k = tid; k < M; k += kSolveThreads) {
    double y[3] = {tk[k], tot[5 + 2 * k], tot[6 + 2 * k]};
    chol3_solve(s_u, y);
    const double d = atan2(-y[2], y[1]);
    dnew[k] = d;
    double sn, c;
    sincos(d, &sn, &c);
    cs[k] = c;
    cs[M + k] = sn;
  }
  __syncthreads();
  // convergence error and the next phase-step matrix A1 = sum_k v_k v_k^T
  double e = 0, acc[5] = {0, 0, 0, 0, 0};
  for (int k = tid; k < M; k += kSolveThreads) {
    const double d = err_mode ? fabs((dnew[k] - dnew[0]) - (dold[k] - dold[0]))
                              : fabs(dnew[k] - dold[k]);
    e = fmax(e, d);
    const double c = cs[k], sn = cs[M + k];
    acc[0] += c; acc[1] += c * c; acc[2] += sn; acc[3] += c * sn; acc[4] += sn * sn;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
#pragma unroll
  for (int i = 0; i < 5; ++i) acc[i] = wave_sum(acc[i]);
  if (lane == 0) {
    s_err[wid] = e;
#pragma unroll
    for (int i = 0; i < 5; ++i) s_red[wid * 5 + i] = acc[i];
  }
  __syncthreads();
  if (tid == 0) {
    double t[5] = {0, 0, 0, 0, 0}, m = 0;
    for (int w = 0; w < kWaves; ++w) {
      m = fmax(m, s_err[w]);
      for (int i = 0; i < 5; ++i) t[i] += s_red[w * 5 + i];
    }
    const double a[6] = {(double)M, t[0], t[2], t[1], t[3], t[4]};
    s_info = chol3(a, s_u);
    const int it = s_iters + 1;
    ctl->iters = it;
    ctl->err = m;
    ctl->status = s_info ? 10 + s_info : 0;
    ctl->done = (s_info != 0) || (timing ? it >= max_iter : !(m > max_err && it < max_iter));
  }
  __syncthreads();
  for (int k = tid; k < M; k += kSolveThreads) {
    delta[k] = dnew[k];
    if (s_info == 0) {
      double b[3] = {1.0, cs[k], cs[M + k]};
      chol3_solve(s_u, b);
      w1g[k] = (float)b[1];
      w2g[k] = (float)b[2];
    }
  }
}

// Final phase (AIA/aia_cudaf.cu:351-356 and :84-108): phi = atan2(-x2, x1)
// with the final weights, phi + delta_0 wrapped into (-pi, pi]; block 0 also
// writes the shifts relative to delta_0, wrapped (aia_cudaf.cu:364-371).
template <typename T, int GS>
__global__ __launch_bounds__(kThreads) void aia_final_kernel(
    const T* __restrict__ f, int M, long long N, int vec, const AiaCtl* __restrict__ ctl,
    const float* __restrict__ w1g, const float* __restrict__ w2g,
    const double* __restrict__ delta, float* __restrict__ phi, float* __restrict__ dout) {
  extern __shared__ unsigned char smem_raw[];
  float* sw1 = reinterpret_cast<float*>(smem_raw);
  const int Mp = padded_frames(M);
  float* sw2 = sw1 + Mp;
  if (ctl->status) return;
  for (int k = threadIdx.x; k < Mp; k += kThreads) {
    sw1[k] = k < M ? w1g[k] : 0.f;
    sw2[k] = k < M ? w2g[k] : 0.f;
  }
  const double d0 = delta[0];
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < M; k += kThreads) {
      double s, c;
      sincos(delta[k] - d0, &s, &c);
      dout[k] = (float)atan2(s, c);
    }
  __syncthreads();
  const long long nq = (N + 3) >> 2;
  for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < nq;
       q += (long long)gridDim.x * kThreads) {
    const long long p0 = q << 2;
    float i0[4], x1[4] = {0, 0, 0, 0}, x2[4] = {0, 0, 0, 0};
    load4<T>(f, p0, N, vec, i0);
    for (int kb = 0; kb < M; kb += GS)
      phase_group<T, GS>(f, kb, M, N, p0, vec, sw1, sw2, i0, x1, x2, nullptr);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p0 + p < N) {
        double t = (double)atan2f(-x2[p], x1[p]) + d0;
        if (t > M_PI) t -= 2.0 * M_PI;
        else if (t <= -M_PI) t += 2.0 * M_PI;
        phi[p0 + p] = (float)t;
      }
    }
  }
}

}  // namespace

// ---- handle -------------------------------------------------------------------
// Device memory (one allocation): control block; shifts (FP64) and phase-step
// weights (FP32, padded to a multiple of 32) contiguous, so one copy sets all
// three from the host; frame sums; pass partials; frame-sum partials; output
// shifts.
struct glsu_aia {
  int M = 0, Mp = 0, W = 0, H = 0;
  long long N = 0;
  hipStream_t s = nullptr;
  int G = 0, Gs = 0, grid_final = 0;
  void* dmem = nullptr;
  AiaCtl* ctl = nullptr;
  double *delta = nullptr, *Tk = nullptr, *sums_part = nullptr;
  float *w1 = nullptr, *w2 = nullptr, *dout = nullptr, *part = nullptr;
  size_t state_bytes = 0;  // delta + w1 + w2
  void* frames = nullptr;  // host-entry staging (lazy)
  size_t frames_bytes = 0;
  float* phi_dev = nullptr;  // host-entry staging (lazy)
  // pinned host: the state block's image, the control block, the output shifts
  void* hmem = nullptr;
  double* h_delta = nullptr;
  float *h_w1 = nullptr, *h_w2 = nullptr;
  AiaCtl* h_ctl = nullptr;
  float* h_dout = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

static int kc_for(int M) { return M <= 8 ? 8 : (M <= 16 ? 16 : (M <= 24 ? 24 : 32)); }

static size_t solve_smem(int M) { return sizeof(double) * (5 + 7 * (size_t)M); }

extern "C" int glsu_aia_create(glsu_aia** out, int nframes, int width, int height,
                               void* hip_stream) {
  if (!out) return afail(GLSU_ERR_ARG, "null output handle");
  *out = nullptr;
  if (nframes < 3 || nframes > kMaxFrames)
    return afail(GLSU_ERR_ARG, "AIA needs 3..%d frames (got %d)", kMaxFrames, nframes);
  if (width < 1 || height < 1) return afail(GLSU_ERR_ARG, "invalid image size %dx%d", width, height);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return afail(GLSU_ERR_DEVICE, "no HIP device visible");
  auto* h = new glsu_aia;
  h->M = nframes;
  h->Mp = padded_frames(nframes);
  h->W = width;
  h->H = height;
  h->N = (long long)width * height;
  h->s = (hipStream_t)hip_stream;
  const long long nq = (h->N + 3) / 4;
  long long G = (nq + kThreads - 1) / kThreads;  // one quad per thread up to 2048 blocks
  if (G > 2048) G = 2048;
  if (G >= 8) G = (G + 7) / 8 * 8;  // block b and its chunk copies on one XCD
  if (G < 1) G = 1;
  h->G = (int)G;
  long long Gs = (nq + 8 * kThreads - 1) / (8 * kThreads);
  h->Gs = (int)(Gs < 1 ? 1 : (Gs > 64 ? 64 : Gs));
  long long gf = (nq + kThreads - 1) / kThreads;
  h->grid_final = (int)(gf > 4096 ? 4096 : gf);
  const int KC = kc_for(nframes), nch = (nframes + KC - 1) / KC;
  const size_t Gtot = (size_t)h->G * nch;
  const size_t M = (size_t)nframes, Mp = (size_t)h->Mp;
  h->state_bytes = sizeof(double) * M + 2 * sizeof(float) * Mp;
  const size_t bytes = 256 + h->state_bytes + 64 +
                       sizeof(double) * (M + M * h->Gs) +
                       sizeof(float) * ((5 + 2 * (size_t)KC) * Gtot + M) + 64;
  auto cleanup = [&](int rc) {
    if (h->dmem) (void)hipFree(h->dmem);
    if (h->hmem) (void)hipHostFree(h->hmem);
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    delete h;
    return rc;
  };
  if (hipMalloc(&h->dmem, bytes) != hipSuccess)
    return cleanup(afail(GLSU_ERR_DEVICE, "hipMalloc of %zu bytes failed", bytes));
  unsigned char* p = (unsigned char*)h->dmem;
  h->ctl = (AiaCtl*)p;
  p += 256;
  h->delta = (double*)p;
  h->w1 = (float*)(p + sizeof(double) * M);
  h->w2 = h->w1 + Mp;
  p += (h->state_bytes + 63) & ~(size_t)63;
  h->Tk = (double*)p; p += sizeof(double) * M;
  h->sums_part = (double*)p; p += sizeof(double) * M * h->Gs;
  h->part = (float*)p; p += sizeof(float) * (5 + 2 * (size_t)KC) * Gtot;
  h->dout = (float*)p;
  if (hipMemset(h->dmem, 0, bytes) != hipSuccess)
    return cleanup(afail(GLSU_ERR_DEVICE, "hipMemset failed"));
  const size_t hbytes = ((h->state_bytes + 63) & ~(size_t)63) + 64 + sizeof(float) * M;
  if (hipHostMalloc(&h->hmem, hbytes, hipHostMallocDefault) != hipSuccess)
    return cleanup(afail(GLSU_ERR_DEVICE, "hipHostMalloc failed"));
  unsigned char* q = (unsigned char*)h->hmem;
  h->h_delta = (double*)q;
  h->h_w1 = (float*)(q + sizeof(double) * M);
  h->h_w2 = h->h_w1 + Mp;
  q += (h->state_bytes + 63) & ~(size_t)63;
  h->h_ctl = (AiaCtl*)q;
  h->h_dout = (float*)(q + 64);
  for (size_t k = 0; k < Mp; ++k) h->h_w1[k] = h->h_w2[k] = 0.f;
  if (hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess)
    return cleanup(afail(GLSU_ERR_DEVICE, "hipEventCreate failed"));
  *out = h;
  return GLSU_OK;
}

extern "C" int glsu_aia_destroy(glsu_aia* h) {
  if (!h) return GLSU_OK;
  if (h->s) (void)hipStreamSynchronize(h->s);
  else (void)hipDeviceSynchronize();
  if (h->dmem) (void)hipFree(h->dmem);
  if (h->frames) (void)hipFree(h->frames);
  if (h->phi_dev) (void)hipFree(h->phi_dev);
  if (h->hmem) (void)hipHostFree(h->hmem);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  delete h;
  return GLSU_OK;
}

extern "C" int glsu_aia_info(glsu_aia* h, int* blocks, int* chunks, int* frames_per_chunk) {
  if (!h) return afail(GLSU_ERR_ARG, "null handle");
  const int KC = kc_for(h->M);
  if (blocks) *blocks = h->G;
  if (chunks) *chunks = (h->M + KC - 1) / KC;
  if (frames_per_chunk) *frames_per_chunk = KC;
  return GLSU_OK;
}

// Shifts and first weights to the device (one copy from the pinned image).
// Returns the phase-step Cholesky status of the initial shifts.
static int stage_initial(glsu_aia* h, const double* d0) {
  for (int k = 0; k < h->M; ++k) h->h_delta[k] = d0[k];
  const int info = phase_weights_host(h->h_delta, h->M, h->h_w1, h->h_w2);
  if (info)
    return afail(GLSU_ERR_NOT_PD,
                 "AIA phase step: the leading minor of order %d is not positive definite; the "
                 "solution could not be computed",
                 info);
  ATRY(hipMemcpyAsync(h->delta, h->h_delta, h->state_bytes, hipMemcpyHostToDevice, h->s));
  return GLSU_OK;
}

template <typename T, int KC>
static hipError_t launch_pass(glsu_aia* h, const void* f, int vec) {
  const int nch = (h->M + KC - 1) / KC;
  hipLaunchKernelGGL((aia_pass_kernel<T, KC>), dim3(h->G * nch), dim3(kThreads),
                     2 * h->Mp * sizeof(float), h->s, (const T*)f, h->M, h->N, vec, h->G, h->ctl,
                     h->w1, h->w2, h->part);
  return hipGetLastError();
}

template <typename T, int KC>
static hipError_t launch_iter(glsu_aia* h, const void* f, int vec, bool first, int max_iter,
                              double max_err, int err_mode, int timing) {
  const int nch = (h->M + KC - 1) / KC;
  hipError_t e = launch_pass<T, KC>(h, f, vec);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((aia_solve_kernel<KC>), dim3(1), dim3(kSolveThreads), solve_smem(h->M), h->s,
                     h->M, h->N, h->G, h->G * nch, h->ctl, h->w1, h->w2, h->delta, h->Tk, h->part,
                     h->sums_part, first ? h->Gs : 0, max_iter, max_err, err_mode, timing);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_iter_any(glsu_aia* h, const void* f, int vec, bool first, int max_iter,
                                  double max_err, int err_mode, int timing) {
  switch (kc_for(h->M)) {
    case 8: return launch_iter<T, 8>(h, f, vec, first, max_iter, max_err, err_mode, timing);
    case 16: return launch_iter<T, 16>(h, f, vec, first, max_iter, max_err, err_mode, timing);
    case 24: return launch_iter<T, 24>(h, f, vec, first, max_iter, max_err, err_mode, timing);
    default: return launch_iter<T, 32>(h, f, vec, first, max_iter, max_err, err_mode, timing);
  }
}

template <typename T>
static hipError_t launch_pass_any(glsu_aia* h, const void* f, int vec) {
  switch (kc_for(h->M)) {
    case 8: return launch_pass<T, 8>(h, f, vec);
    case 16: return launch_pass<T, 16>(h, f, vec);
    case 24: return launch_pass<T, 24>(h, f, vec);
    default: return launch_pass<T, 32>(h, f, vec);
  }
}

template <typename T>
static hipError_t launch_setup(glsu_aia* h, const void* f, int vec, int max_iter, double max_err,
                               int timing) {
  hipLaunchKernelGGL((aia_frame_sums_kernel<T>), dim3(h->Gs, h->M), dim3(kThreads), 0, h->s,
                     (const T*)f, h->N, vec, h->sums_part, h->ctl, max_iter, max_err, timing);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_final(glsu_aia* h, const void* f, int vec, float* phi) {
  const int GS = sizeof(T) > 1 ? 8 : (h->M <= 8 ? 8 : h->M <= 16 ? 16 : h->M <= 24 ? 24 : 32);
  const size_t sm = 2 * h->Mp * sizeof(float);
  const dim3 g(h->grid_final), b(kThreads);
#define GLSU_AIA_FINAL(gs)                                                                      \
  hipLaunchKernelGGL((aia_final_kernel<T, gs>), g, b, sm, h->s, (const T*)f, h->M, h->N, vec,  \
                     h->ctl, h->w1, h->w2, h->delta, phi, h->dout)
  switch (GS) {
    case 8: GLSU_AIA_FINAL(8); break;
    case 16: GLSU_AIA_FINAL(16); break;
    case 24: GLSU_AIA_FINAL(24); break;
    default: GLSU_AIA_FINAL(32); break;
  }
#undef GLSU_AIA_FINAL
  return hipGetLastError();
}

template <typename T>
static hipError_t enqueue_typed(glsu_aia* h, const void* f, int vec, int max_iter, double max_err,
                                int err_mode, float* phi) {
  hipError_t e = launch_setup<T>(h, f, vec, max_iter, max_err, 0);
  if (e != hipSuccess) return e;
  for (int it = 0; it < max_iter; ++it) {
    if (it > 0 && it % kLaunchGroup == 0) {
      // poll the convergence flag between groups of launches
      e = hipMemcpyAsync(h->h_ctl, h->ctl, sizeof(AiaCtl), hipMemcpyDeviceToHost, h->s);
      if (e == hipSuccess) e = hipStreamSynchronize(h->s);
      if (e != hipSuccess) return e;
      if (h->h_ctl->done) break;
    }
    e = launch_iter_any<T>(h, f, vec, it == 0, max_iter, max_err, err_mode, 0);
    if (e != hipSuccess) return e;
  }
  return launch_final<T>(h, f, vec, phi);
}

static int frame_vec(const glsu_aia* h, const void* f, int dtype) {
  const size_t es = dtype == GLSU_U8 ? 1 : (dtype == GLSU_F32 ? 4 : 8);
  return (h->N % 4 == 0) && ((uintptr_t)f % (4 * es) == 0);
}

static int aia_run_dev(glsu_aia* h, const void* d_frames, int dtype, float* deltas, int use_given,
                       uint64_t seed, int max_iter, double max_err, int err_mode, float* d_phi,
                       int* iters, double* err, double* ms) {
  if (!h) return afail(GLSU_ERR_ARG, "null handle");
  if (!d_frames || !d_phi || !deltas)
    return afail(GLSU_ERR_ARG, "frames, phi and deltas must not be null");
  if (dtype != GLSU_U8 && dtype != GLSU_F32 && dtype != GLSU_F64)
    return afail(GLSU_ERR_ARG, "unknown frame dtype %d", dtype);
  if (err_mode != 0 && err_mode != 1) return afail(GLSU_ERR_ARG, "err_mode must be 0 or 1");
  if (max_iter < 0) return afail(GLSU_ERR_ARG, "max_iter must be >= 0");
  const int M = h->M;
  // initial shifts: given, or delta_0 = 0 and N(0,1) draws (AIA/aia_cpuf.cpp:31-45,
  // which seeds from the clock; here the seed is an argument)
  std::vector<double> d0(M);
  if (use_given) {
    for (int k = 0; k < M; ++k) d0[k] = (double)deltas[k];
  } else {
    std::mt19937_64 gen(seed);
    std::normal_distribution<float> dist(0.0f, 1.0f);
    d0[0] = 0.0;
    for (int k = 1; k < M; ++k) d0[k] = (double)dist(gen);
  }
  const int vec = frame_vec(h, d_frames, dtype);
  int rc = stage_initial(h, d0.data());
  if (rc != GLSU_OK) return rc;
  ATRY(hipEventRecord(h->e0, h->s));
  hipError_t e;
  switch (dtype) {
    case GLSU_U8: e = enqueue_typed<uint8_t>(h, d_frames, vec, max_iter, max_err, err_mode, d_phi); break;
    case GLSU_F32: e = enqueue_typed<float>(h, d_frames, vec, max_iter, max_err, err_mode, d_phi); break;
    default: e = enqueue_typed<double>(h, d_frames, vec, max_iter, max_err, err_mode, d_phi); break;
  }
  if (e != hipSuccess) return afail(GLSU_ERR_DEVICE, "AIA launch failed: %s", hipGetErrorString(e));
  ATRY(hipEventRecord(h->e1, h->s));
  ATRY(hipMemcpyAsync(h->h_ctl, h->ctl, sizeof(AiaCtl), hipMemcpyDeviceToHost, h->s));
  ATRY(hipMemcpyAsync(h->h_dout, h->dout, sizeof(float) * M, hipMemcpyDeviceToHost, h->s));
  ATRY(hipStreamSynchronize(h->s));
  const AiaCtl c = *h->h_ctl;
  if (c.status) {
    const int step = c.status / 10, k = c.status % 10;
    return afail(GLSU_ERR_NOT_PD,
                 "AIA %s step: the leading minor of order %d is not positive definite; the "
                 "solution could not be computed",
                 step == 1 ? "phase" : "shift", k);
  }
  for (int k = 0; k < M; ++k) deltas[k] = h->h_dout[k];
  if (iters) *iters = c.iters;
  if (err) *err = c.err;
  if (ms) {
    float t = 0;
    ATRY(hipEventElapsedTime(&t, h->e0, h->e1));
    *ms = t;
  }
  return GLSU_OK;
}

extern "C" int glsu_aia_run(glsu_aia* h, const void* d_frames, int frame_dtype, float* deltas,
                            int use_given, uint64_t seed, int max_iter, double max_err,
                            int err_mode, float* d_phi, int* iters, double* err, double* ms) {
  return aia_run_dev(h, d_frames, frame_dtype, deltas, use_given, seed, max_iter, max_err,
                     err_mode, d_phi, iters, err, ms);
}

extern "C" int glsu_aia_run_host(glsu_aia* h, const void* const* h_frames, int frame_dtype,
                                 float* deltas, int use_given, uint64_t seed, int max_iter,
                                 double max_err, int err_mode, float* h_phi, int* iters,
                                 double* err, double* ms) {
  if (!h) return afail(GLSU_ERR_ARG, "null handle");
  if (!h_frames || !h_phi) return afail(GLSU_ERR_ARG, "frames and phi must not be null");
  if (frame_dtype != GLSU_U8 && frame_dtype != GLSU_F32 && frame_dtype != GLSU_F64)
    return afail(GLSU_ERR_ARG, "unknown frame dtype %d", frame_dtype);
  const size_t es = frame_dtype == GLSU_U8 ? 1 : (frame_dtype == GLSU_F32 ? 4 : 8);
  const size_t fb = es * (size_t)h->N;
  if (h->frames_bytes < fb * h->M) {
    if (h->frames) (void)hipFree(h->frames);
    h->frames = nullptr;
    h->frames_bytes = 0;
    ATRY(hipMalloc(&h->frames, fb * h->M));
    h->frames_bytes = fb * h->M;
  }
  if (!h->phi_dev) ATRY(hipMalloc((void**)&h->phi_dev, sizeof(float) * (size_t)h->N));
  for (int k = 0; k < h->M; ++k) {
    if (!h_frames[k]) return afail(GLSU_ERR_ARG, "frame %d is null", k);
    ATRY(hipMemcpyAsync((unsigned char*)h->frames + fb * k, h_frames[k], fb,
                        hipMemcpyHostToDevice, h->s));
  }
  const int rc = aia_run_dev(h, h->frames, frame_dtype, deltas, use_given, seed, max_iter,
                             max_err, err_mode, h->phi_dev, iters, err, ms);
  if (rc != GLSU_OK) return rc;
  ATRY(hipMemcpyAsync(h_phi, h->phi_dev, sizeof(float) * (size_t)h->N, hipMemcpyDeviceToHost,
                      h->s));
  ATRY(hipStreamSynchronize(h->s));
  return GLSU_OK;
}

// Average time of one iteration (stage 0: pass + solve launches) or of the
// pass kernel alone (stage 1) over reps repetitions, with the convergence test
// off (every iteration runs; the shift estimates update every iteration), timed
// with events on the handle's stream.  No reference counterpart.
template <typename T>
static int time_iter_typed(glsu_aia* h, const void* f, int vec, int stage, int reps, double* ms) {
  const int total = stage == 0 ? reps + 1 : 2;  // stage 1: the warm-up leaves done = 0
  hipError_t e = launch_setup<T>(h, f, vec, total, 1.0, 1);
  if (e == hipSuccess) e = launch_iter_any<T>(h, f, vec, true, total, 1.0, 0, 1);  // warm-up
  if (e != hipSuccess) return afail(GLSU_ERR_DEVICE, "AIA launch failed: %s", hipGetErrorString(e));
  ATRY(hipEventRecord(h->e0, h->s));
  for (int r = 0; r < reps; ++r) {
    e = stage == 0 ? launch_iter_any<T>(h, f, vec, false, total, 1.0, 0, 1)
                   : launch_pass_any<T>(h, f, vec);
    if (e != hipSuccess) return afail(GLSU_ERR_DEVICE, "AIA launch failed: %s", hipGetErrorString(e));
  }
  ATRY(hipEventRecord(h->e1, h->s));
  ATRY(hipMemcpyAsync(h->h_ctl, h->ctl, sizeof(AiaCtl), hipMemcpyDeviceToHost, h->s));
  ATRY(hipStreamSynchronize(h->s));
  if (h->h_ctl->status) return afail(GLSU_ERR_NOT_PD, "AIA normal matrix not positive definite");
  if (stage == 0 && h->h_ctl->iters != total)
    return afail(GLSU_ERR_DEVICE, "AIA timing: %d of %d iterations ran", h->h_ctl->iters, total);
  float t = 0;
  ATRY(hipEventElapsedTime(&t, h->e0, h->e1));
  *ms = t / reps;
  return GLSU_OK;
}

extern "C" int glsu_aia_time_iter(glsu_aia* h, const void* d_frames, int frame_dtype,
                                  const float* deltas, int stage, int reps, double* ms) {
  if (!h || !d_frames || !deltas || !ms || reps < 1 || (stage != 0 && stage != 1))
    return afail(GLSU_ERR_ARG, "invalid argument");
  if (frame_dtype != GLSU_U8 && frame_dtype != GLSU_F32 && frame_dtype != GLSU_F64)
    return afail(GLSU_ERR_ARG, "unknown frame dtype %d", frame_dtype);
  std::vector<double> d0(h->M);
  for (int k = 0; k < h->M; ++k) d0[k] = (double)deltas[k];
  const int rc = stage_initial(h, d0.data());
  if (rc != GLSU_OK) return rc;
  const int vec = frame_vec(h, d_frames, frame_dtype);
  switch (frame_dtype) {
    case GLSU_U8: return time_iter_typed<uint8_t>(h, d_frames, vec, stage, reps, ms);
    case GLSU_F32: return time_iter_typed<float>(h, d_frames, vec, stage, reps, ms);
    default: return time_iter_typed<double>(h, d_frames, vec, stage, reps, ms);
  }
}

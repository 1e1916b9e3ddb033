// wff_kernels.h -- WFF2 kernel templates (included by the wff_*.hip units).
// WFF2 on gfx950: separable windowed-Fourier filtering.
//
// Reference semantics (WFT/WFT2_CPU.cpp:190-390):
//   F = wxi*wyi/(4 pi^2) * sum_{wy,wx} w_p * T( w_p * f ),
//   w_p(y,x) = g(y,x) exp(j(wx x + wy y)),  T(u) = u if |u| >= thr else 0,
// with * the linear "same" convolution and T applied on the image only.
//
// The window is separable, g = gy(y) gx(x), so w_p = [gy e^{j wy y}] [gx e^{j wx x}]
// and each 2-D convolution is a convolution along x followed by one along y.
// The pair loop is reorganised as
//   A_wx   = f (*)_x hx_wx                         (rows:    1 FFT + nwx IFFTs per row)
//   D_wx   = sum_wy hy_wy (*)_y T(hy_wy (*)_y A_wx) (columns: 2 + 2 nwy FFTs per column)
//   F      = c * sum_wx hx_wx (*)_x D_wx            (rows:    nwx FFTs + 1 IFFT per row)
// All convolutions are circular of length Nx >= W + sx and Ny >= H + sy with a
// centred kernel, which equals the linear convolution on the image exactly.
// The whole wy loop of a column runs out of VGPRs and LDS: the column
// spectrum, its per-pair products and the running sum never leave the CU.
#pragma once
#include "fft_device.h"
#include "glsu_internal.h"

#include <algorithm>
#include <cmath>

namespace glsu {
// ------------------------------------------------------------ plan table
#define GLSU_PLANS(X)                    \
  X(0, 64, 4, 4, 4, 4)                   \
  X(1, 64, 8, 8, 8)                      \
  X(2, 64, 12, 4, 4, 4)                  \
  X(3, 64, 16, 4, 4, 4)                  \
  X(4, 64, 20, 4, 4, 4)                  \
  X(5, 64, 24, 8, 8)                     \
  X(6, 64, 28, 4, 4, 4)                  \
  X(7, 64, 32, 8, 8)                     \
  X(8, 64, 36, 4, 4, 4)                  \
  X(9, 64, 40, 8, 8)                     \
  X(10, 16, 8, 4, 4)                     \
  X(11, 16, 12, 4, 4)                    \
  X(12, 16, 20, 4, 4)                    \
  X(13, 32, 12, 4, 4, 2)                 \
  X(14, 32, 20, 4, 4, 2)

// The same table as constants (no allocation at any call).
struct PlanRow {
  int id, T, E, N;
};
#define GLSU_PLAN_ROW(id, T, E, ...) {id, T, E, T * E},
inline constexpr PlanRow kPlanRows[] = {GLSU_PLANS(GLSU_PLAN_ROW)};
#undef GLSU_PLAN_ROW
inline const PlanRow* plan_row(int id) {
  for (const PlanRow& r : kPlanRows)
    if (r.id == id) return &r;
  return nullptr;
}

// Live registers of a transform: when the image extent n along the axis fits
// in the first EV registers of every lane (n <= EV * T), the last E - EV
// registers only ever hold zero inputs and discarded outputs; kernels compiled
// for that EV treat them as such at compile time, and the compiler drops the
// arithmetic that feeds or reads only them (the 1280-point plan at 1024
// pixels: 16 of 20 registers live).  Used by the row-forward and column
// kernels, single and batched alike (batch classes are split by whether their
// images fit, so both forms round the same way); the row-inverse kernel
// measured slower with it and keeps all E.
template <class P>
inline constexpr int kLiveRegs = live_regs(P::E);
template <class P>
inline bool live_regs_fit(int n) {
  return n <= kLiveRegs<P> * P::T;
}

__device__ __forceinline__ size_t tile_index(int y, int x, int H) {
  return ((size_t)(x / kTileCols) * H + y) * kTileCols + (x % kTileCols);
}

// ---------------------------------------------------------- row forward
// One workgroup per row, RW waves (one transform slot each, T == 64).
// Wave 0 computes Zr = FFT_x(f row) and shares it through LDS; wave w then
// produces A_wx = IFFT_x(Zr * Hx_wx) for wx = w, w + RW, ...
template <class P, int EV = P::E>
__device__ __forceinline__ void rows_fwd_body(const float2* __restrict__ frow,
                                              float2* __restrict__ A,
                                              const float* __restrict__ Hx,
                                              const float2* __restrict__ tw_g, int W, int H,
                                              int nwx, size_t plane, int y, int lane,
                                              float2* smem) {
  static_assert(P::T == 64, "row kernels use one wavefront per transform");
  constexpr int T = P::T, E = P::E, N = P::N;
  const int RW = blockDim.x / T;
  const int wave = threadIdx.x / T;
  float2* buf = smem + wave * P::LDS;
  FftLane<P> ctx;
  ctx.init(lane, tw_g);
  float2 z[E];
  if (wave == 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int x = lane + T * j;
      z[j] = j < EV && x < W ? frow[x] : make_float2(0.f, 0.f);
    }
    fft_fwd<P>(z, buf, ctx);
#pragma unroll
    for (int j = 0; j < E; ++j) buf[j * T + lane] = z[j];  // share the spectrum
  }
  __syncthreads();
  if (wave != 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) z[j] = smem[j * T + lane];
  }
  __syncthreads();  // wave 0 reuses its buffer below
  // long transforms (one wavefront per SIMD): the next wx's spectrum is read
  // before this wx's plane stores (the wait for a load also waits for every
  // store issued before it); the extra registers spill to AGPRs
  constexpr bool kPrefetch = E > 24;
  float h[E];  // used only when kPrefetch
  auto load_h = [&](int k) {
    const float* hx = Hx + (size_t)k * N + lane;  // lane-major: one 256-byte row per j
#pragma unroll
    for (int j = 0; j < E; ++j) h[j] = hx[j * T];
  };
  if constexpr (kPrefetch)
    if (wave < nwx) load_h(wave);
  for (int k = wave; k < nwx; k += RW) {
    float2 v[E];
    if constexpr (kPrefetch) {
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = rmul(z[j], h[j]);
    } else {
      const float* hx = Hx + (size_t)k * N + lane;
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = rmul(z[j], hx[j * T]);
    }
    fft_inv<P>(v, buf, ctx);
    if constexpr (kPrefetch)
      if (k + RW < nwx) load_h(k + RW);
    float2* out = A + plane * k;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int x = lane + T * j;
      if (j < EV && x < W) out[tile_index(y, x, H)] = v[j];
    }
  }
}

template <class P, int EV = P::E>
__global__ __launch_bounds__(64 * kRowWavesMax) void wff_rows_fwd_kernel(
    const float2* __restrict__ f, int ld, float2* __restrict__ A, const float* __restrict__ Hx,
    const float2* __restrict__ tw_g, int W, int H, int nwx, size_t plane,
    const int* __restrict__ skip) {
  extern __shared__ float2 smem[];
  if (skip && *skip) return;  // device-side decision (learned path): not needed
  const int y = xcd_remap(blockIdx.x, gridDim.x);  // uniform per workgroup
  if (y >= H) return;
  rows_fwd_body<P, EV>(f + (size_t)y * ld, A, Hx, tw_g, W, H, nwx, plane, y, threadIdx.x % P::T,
                   smem);
}

// -------------------------------------------------------------- columns
// One transform slot per (wx, column).  Ahat = FFT_y(A col); for each wy:
// S = IFFT_y(Ahat*Hy); T = threshold(S) on image rows; acc += FFT_y(T)*Hy.
// D col = IFFT_y(acc).  `ttab` (transition twiddles in LDS) is filled by the
// caller before the first item.  HOLD: keep each wy spectrum in registers
// across both of its products, else read it again for the second one.
// ZSKIP: how a skipped transform rejoins the loop (below).
template <class P, bool HOLD = true, bool ZSKIP = true, int EV = P::E>
__device__ __forceinline__ void cols_body(const float2* __restrict__ A, float2* __restrict__ D,
                                          const float* __restrict__ Hy,
                                          const float2* __restrict__ tw_g, float thr, int W,
                                          int H, int nwx, int nwy, size_t plane, int item,
                                          int lane, const float2* ttab, float2* buf,
                                          unsigned int* nskip = nullptr) {
  constexpr int T = P::T, E = P::E, N = P::N;
  using Ctx = FftLane<P, true>;
  const int k = item / W, x = item % W;
  const float thr2 = thr * thr;
  Ctx ctx;
  ctx.init(lane, tw_g, ttab);
  const float2* colA = A + plane * k;
  float2 ah[E], acc[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int y = lane + T * j;
    ah[j] = j < EV && y < H ? colA[tile_index(y, x, H)] : make_float2(0.f, 0.f);
    acc[j] = make_float2(0.f, 0.f);
  }
  fft_fwd<P>(ah, buf, ctx);
  // the wy spectrum (real: E registers) is read once and held across both of
  // its products; at one wavefront per SIMD (E > 24) nothing else would hide
  // a second L2 round trip, and at two the held copy still fits under 256 VGPRs
  // in the single-image kernel (the persistent batched kernel's own state
  // leaves no room for it there: it reads the spectrum again)
  for (int q = 0; q < nwy; ++q) {
    const float* hy = Hy + (size_t)q * N + lane;  // lane-major: one T-float run per j
    float h[E];
    float2 v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if constexpr (HOLD) {
        h[j] = hy[j * T];
        v[j] = rmul(ah[j], h[j]);
      } else {
        v[j] = rmul(ah[j], hy[j * T]);
      }
    }
    fft_inv<P>(v, buf, ctx);
    bool live = false;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int y = lane + T * j;
      const float m2 = v[j].x * v[j].x + v[j].y * v[j].y;
      const bool keep = j < EV && y < H && m2 >= thr2;
      if (!keep) v[j] = make_float2(0.f, 0.f);
      live |= keep;
    }
    // a column whose thresholded signal is zero everywhere adds exactly
    // nothing: skip its transform when no slot of the wavefront has a
    // surviving sample (wave-uniform branch).  ZSKIP: both branches define the
    // same registers (zeros here) and the sum below runs on both, adding exact
    // zeros after a skip; with a `continue`, the two paths into the loop latch
    // hold the running sum in different registers and every iteration pays 2E
    // register copies.  In the persistent batched kernel the zero-sum form
    // costs registers (an occupancy step at E = 20), so it keeps the `continue`.
    if (__builtin_amdgcn_ballot_w64(live) == 0) {
      if (nskip && lane == 0) atomicAdd(nskip, 1u);  // diagnostics: transforms skipped
      if constexpr (!ZSKIP) continue;
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = make_float2(0.f, 0.f);
    } else {
      fft_fwd<P>(v, buf, ctx);
      if constexpr (!HOLD) {
        // the opaque copy of the lane offset defeats load CSE with the first read
        int l2 = lane;
        __asm__ volatile("" : "+v"(l2));
        const float* hy2 = Hy + (size_t)q * N + l2;
#pragma unroll
        for (int j = 0; j < E; ++j) h[j] = hy2[j * T];
      }
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
      acc[j].x += v[j].x * h[j];
      acc[j].y += v[j].y * h[j];
    }
  }
  fft_inv<P>(acc, buf, ctx);
  float2* colD = D + plane * k;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int y = lane + T * j;
    if (j < EV && y < H) colD[tile_index(y, x, H)] = acc[j];
  }
}

template <class P, int WAVES, int EV = P::E>
__global__ __launch_bounds__(64 * WAVES) void wff_cols_kernel(
    const float2* __restrict__ A, float2* __restrict__ D, const float* __restrict__ Hy,
    const float2* __restrict__ tw_g, const float* __restrict__ thr_p, int W, int H, int nwx,
    int nwy, size_t plane, const int* __restrict__ skip, unsigned int* __restrict__ nskip) {
  constexpr int T = P::T, E = P::E;
  extern __shared__ float2 smem[];
  if (skip && *skip) return;  // device-side decision (learned path): not needed
  float2* ttab = smem;
  FftLane<P, true>::fill_ttab(ttab, tw_g);
  __syncthreads();
  const int slot = threadIdx.x / T;
  float2* buf = smem + E * T + slot * P::LDS;
  // block b: plane k = b % nwx, columns [(b / nwx) * ipb, +ipb).  Planes cycle
  // fastest so that every XCD's contiguous block range holds an even mix of
  // wx: columns whose thresholded signal is zero (skipped transforms) cluster
  // by wx, and the kernel lasts as long as its busiest XCD.
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int x = (b / nwx) * (blockDim.x / T) + slot;
  if (x >= W) return;
  cols_body<P, true, true, EV>(A, D, Hy, tw_g, *thr_p, W, H, nwx, nwy, plane, (b % nwx) * W + x,
               threadIdx.x % T, ttab, buf, nskip);
}

// ---------------------------------------------------------- row inverse
// One workgroup per row, RW waves.  Wave w accumulates FFT_x(D_wx row) * Hx_wx
// over wx = w, w + RW, ...; the partial spectra are summed through LDS and
// wave 0 computes F = scale * IFFT_x(sum).
// MODE 0: write F.  MODE 1: DPRA phase update (DPRA/dpra_cpu.cpp:244-251):
//   t = angle(F); phicurr = wrap(phiref + t); dphi = dphiref + t (two-float sum).
template <class P, int MODE>
__device__ __forceinline__ void rows_inv_body(const float2* __restrict__ D,
                                              const float* __restrict__ Hx,
                                              const float2* __restrict__ tw_g, int W, int H,
                                              int nwx, size_t plane, float scale,
                                              float2* __restrict__ filtered, const UpdateArgs& u,
                                              int y, int lane, float2* smem) {
  static_assert(P::T == 64, "row kernels use one wavefront per transform");
  constexpr int T = P::T, E = P::E, N = P::N;
  const int RW = blockDim.x / T;
  const int wave = threadIdx.x / T;
  float2* buf = smem + wave * P::LDS;
  FftLane<P> ctx;
  ctx.init(lane, tw_g);
  float2 acc[E];
#pragma unroll
  for (int j = 0; j < E; ++j) acc[j] = make_float2(0.f, 0.f);
  if constexpr (E > 24) {
    // long transforms (one wavefront per SIMD) read the next wx's row during
    // this one's transform; the extra registers spill to AGPRs
    auto load_row = [&](float2(&r)[E], int k) {
      const float2* Dk = D + plane * k;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int x = lane + T * j;
        r[j] = x < W ? Dk[tile_index(y, x, H)] : make_float2(0.f, 0.f);
      }
    };
    float2 nxt[E];
    if (wave < nwx) load_row(nxt, wave);
    for (int k = wave; k < nwx; k += RW) {
      float2 v[E];
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = nxt[j];
      if (k + RW < nwx) load_row(nxt, k + RW);
      fft_fwd<P>(v, buf, ctx);
      const float* hx = Hx + (size_t)k * N + lane;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const float h = hx[j * T];
        acc[j].x += v[j].x * h;
        acc[j].y += v[j].y * h;
      }
    }
  } else {
    for (int k = wave; k < nwx; k += RW) {
      const float2* Dk = D + plane * k;
      float2 v[E];
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int x = lane + T * j;
        v[j] = x < W ? Dk[tile_index(y, x, H)] : make_float2(0.f, 0.f);
      }
      fft_fwd<P>(v, buf, ctx);
      const float* hx = Hx + (size_t)k * N + lane;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const float h = hx[j * T];
        acc[j].x += v[j].x * h;
        acc[j].y += v[j].y * h;
      }
    }
  }
  if (RW > 1) {
    if (wave != 0) {
#pragma unroll
      for (int j = 0; j < E; ++j) buf[j * T + lane] = acc[j];
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 1; w < RW; ++w) {
      const float2* o = smem + w * P::LDS;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const float2 a = o[j * T + lane];
        acc[j].x += a.x;
        acc[j].y += a.y;
      }
    }
  }
  fft_inv<P>(acc, buf, ctx);
  if constexpr (MODE == 0) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int x = lane + T * j;
      if (x >= W) continue;
      filtered[(size_t)y * W + x] = make_float2(acc[j].x * scale, acc[j].y * scale);
    }
  } else {
    // every input of the update is loaded (clamped columns, always in the
    // image) before the first output is stored
    float pr[E], hi[E], lo[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const size_t id = (size_t)y * W + min(lane + T * j, W - 1);
      pr[j] = u.phiref[id];
      hi[j] = u.dref_hi[id];
      lo[j] = u.dref_lo[id];
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int x = lane + T * j;
      if (x >= W) continue;
      const size_t id = (size_t)y * W + x;
      const float2 F = make_float2(acc[j].x * scale, acc[j].y * scale);
      if (u.filtered) u.filtered[id] = F;
      phase_store(u, id, phase_step(pr[j], hi[j], lo[j], atan2f(F.y, F.x)));
    }
  }
}

template <class P, int MODE>
__global__ __launch_bounds__(64 * kRowWavesMax) void wff_rows_inv_kernel(
    const float2* __restrict__ D, const float* __restrict__ Hx, const float2* __restrict__ tw_g,
    int W, int H, int nwx, size_t plane, float scale, float2* __restrict__ filtered,
    UpdateArgs u, const int* __restrict__ skip) {
  extern __shared__ float2 smem[];
  if (skip && *skip) return;  // device-side decision (learned path): not needed
  const int y = xcd_remap(blockIdx.x, gridDim.x);
  if (y >= H) return;
  rows_inv_body<P, MODE>(D, Hx, tw_g, W, H, nwx, plane, scale, filtered, u, y,
                         threadIdx.x % P::T, smem);
}

// ------------------------------------------------- batched (many images)
// Persistent kernels over a device table of images (BatchSeg) that share one
// FFT plan pair: each workgroup walks virtual work items v = blockIdx.x,
// blockIdx.x + gridDim.x, ... (XCD-contiguous order), finds the image that
// owns the item by binary search on the table's offsets, and runs the same
// per-item body as the single-image kernels, so every image's result is
// bit-identical to a single-image run with the same plan and launch shape.

// The lane index, recomputed behind an empty asm for every item: per-lane
// twiddles and addresses derived from it are then rebuilt per item instead of
// being hoisted out of the persistent loop, where they would stay live in
// VGPRs across the whole body and halve occupancy.
template <class P>
__device__ __forceinline__ int opaque_lane() {
  int lane = threadIdx.x % P::T;
  __asm__ volatile("" : "+v"(lane));
  return lane;
}

// index of the last table entry whose offset (field `F`) is <= key
template <int BatchSeg::*F>
__device__ __forceinline__ int seg_find(const BatchSeg* __restrict__ tab, int n, int key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].*F <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class P, int EV = P::E>
__global__ __launch_bounds__(64 * kRowWavesMax) void wff_rows_fwd_batch_kernel(
    const BatchSeg* __restrict__ tab, int nseg, int nrows, const float* __restrict__ Hx,
    const float2* __restrict__ tw_g, int nwx) {
  extern __shared__ float2 smem[];
  for (int v = blockIdx.x; v < nrows; v += gridDim.x) {
    const int r = xcd_remap(v, nrows);
    const BatchSeg& s = tab[seg_find<&BatchSeg::row0>(tab, nseg, r)];
    __syncthreads();  // LDS of the previous row is free
    rows_fwd_body<P, EV>(s.z + (size_t)(r - s.row0) * s.W, s.A, Hx, tw_g, s.W, s.H, nwx, s.plane,
                     r - s.row0, opaque_lane<P>(), smem);
  }
}

template <class P, int WAVES, int EV = P::E>
__global__ __launch_bounds__(64 * WAVES) void wff_cols_batch_kernel(
    const BatchSeg* __restrict__ tab, int nseg, int nblk, const float* __restrict__ Hy,
    const float2* __restrict__ tw_g, int nwx, int nwy) {
  constexpr int T = P::T, E = P::E;
  extern __shared__ float2 smem[];
  float2* ttab = smem;
  FftLane<P, true>::fill_ttab(ttab, tw_g);
  __syncthreads();
  const int slot = threadIdx.x / T;
  float2* buf = smem + E * T + slot * P::LDS;
  for (int v = blockIdx.x; v < nblk; v += gridDim.x) {
    const int b = xcd_remap(v, nblk);
    const BatchSeg& s = tab[seg_find<&BatchSeg::col0>(tab, nseg, b)];
    const int lb = b - s.col0;  // plane lb % nwx, columns [(lb / nwx) * ipb, +ipb)
    const int x = (lb / nwx) * (blockDim.x / T) + slot;
    if (x < s.W)
      cols_body<P, (P::E > 24), false, EV>(s.A, s.D, Hy, tw_g, *s.thr, s.W, s.H, nwx, nwy,
                                           s.plane, (lb % nwx) * s.W + x, opaque_lane<P>(),
                                           ttab, buf);
  }
}

// number of reference-phase swaps before frame t (swap after frame i when
// i % rr == 0, DPRA/dpra_cpu.cpp:84-87); its parity says which buffer of each
// double buffer holds the reference
__device__ __forceinline__ int swap_parity(int t, int rr) {
  return (t == 0 ? 0 : (t - 1) / rr + 1) & 1;
}

// (two resident workgroups per CU keep the short plans at 2 waves/SIMD; the
// plans of 24 points per lane and more need the whole register file of one
// wave, or they spill)
template <class P>
__global__ __launch_bounds__(64 * kRowWavesMax, P::E >= 24 ? 1 : 2) void wff_rows_inv_batch_kernel(
    const BatchSeg* __restrict__ tab, int nseg, int nrows, const float* __restrict__ Hx,
    const float2* __restrict__ tw_g, int nwx, float scale, const int* __restrict__ frame) {
  extern __shared__ float2 smem[];
  const int t = *frame;
  for (int v = blockIdx.x; v < nrows; v += gridDim.x) {
    const int r = xcd_remap(v, nrows);
    const BatchSeg& s = tab[seg_find<&BatchSeg::row0>(tab, nseg, r)];
    const int c = swap_parity(t, s.rr);
    UpdateArgs u{s.phi[c], s.dhi[c], s.dlo[c], s.phi[c ^ 1], s.dhi[c ^ 1], s.dlo[c ^ 1],
                 s.dphi_out ? s.dphi_out + (size_t)t * s.W * s.H : nullptr, nullptr};
    __syncthreads();
    rows_inv_body<P, 1>(s.D, Hx, tw_g, s.W, s.H, nwx, s.plane, scale, nullptr, u, r - s.row0,
                        opaque_lane<P>(), smem);
  }
}

// ============================================================ host side

template <class P>
static constexpr int plan_lds_elems() { return P::LDS; }

// FftPlan::LDS of every compiled plan, by id
static int plan_lds(int id) {
  switch (id) {
#define GLSU_LDS_CASE(pid, T, E, ...) \
  case pid:                           \
    return FftPlan<T, E, __VA_ARGS__>::LDS;
    GLSU_PLANS(GLSU_LDS_CASE)
#undef GLSU_LDS_CASE
    default:
      return 0;
  }
}

[[maybe_unused]] static int plan_smem_bytes(int id, int waves, bool ttab = false) {
  // per transform: the FftPlan::LDS exchange buffer, plus the shared E x T
  // transition-twiddle table when the kernel keeps it in LDS
  const PlanRow* p = plan_row(id);
  if (!p) return 0;
  return (int)sizeof(float2) * ((waves * 64 / p->T) * plan_lds(id) + (ttab ? p->E * p->T : 0));
}

template <class P>
static int items_per_block(int waves) {
  return waves * 64 / P::T;
}

#define GLSU_DISPATCH_CASE(id, T, E, ...) \
  case id: {                              \
    using P = FftPlan<T, E, __VA_ARGS__>; \
    BODY;                                 \
  } break;

}  // namespace glsu

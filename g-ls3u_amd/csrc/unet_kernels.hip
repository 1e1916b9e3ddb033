// unet_kernels.hip -- kernels of the learned WFF2 surrogate (a 4-level U-Net)
// on gfx950: 3x3 convolutions as implicit GEMMs on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) for the forward pass, the
// input gradient and the weight gradient, plus the small VALU kernels around
// them (2-channel input layer, 1x1 output layer with the loss, max pooling,
// upsampling adjoint, Adam).
//
// Layouts: activations NHWC bf16 (channels innermost, counts multiples of 32
// except the 2-channel input/output); weights [Cout][9][Cin] bf16 for the
// forward pass (K = tap*Cin + ci contiguous per output channel) and
// [Cin][9][Cout] with flipped taps for the input gradient; FP32 master
// weights, gradients and Adam moments in the [Cout][3][3][Cin] order.
#include "unet_internal.h"

#include <algorithm>

namespace glsu {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.f;
  return z;
}

// 8 channels [c, c+8) of pixel (n, y, x) of the (virtual) concatenation
// cat(X1, up2(X2)); zero outside the image.  c is a multiple of 8 and a
// 32-channel chunk never straddles the two sources (C1 % 32 == 0).
__device__ __forceinline__ bf16x8 load8(const __bf16* __restrict__ X1, int C1,
                                        const __bf16* __restrict__ X2, int C2, int n, int y,
                                        int x, int H, int W, int c) {
  if (y < 0 || y >= H || x < 0 || x >= W) return zero8();
  if (c < C1) return *(const bf16x8*)(X1 + (((size_t)n * H + y) * W + x) * C1 + c);
  const int H2 = H >> 1, W2 = W >> 1;
  return *(const bf16x8*)(X2 + (((size_t)n * H2 + (y >> 1)) * W2 + (x >> 1)) * C2 + (c - C1));
}

// The same value as load8 without a branch around the load: the pixel is
// clamped into the image, loaded unconditionally and replaced by zero when it
// was outside, so that a thread's loads can all be in flight together.
__device__ __forceinline__ bf16x8 load8_clamped(const __bf16* __restrict__ X1, int C1,
                                                const __bf16* __restrict__ X2, int C2, int n,
                                                int y, int x, int H, int W, int c) {
  const bool in = y >= 0 && y < H && x >= 0 && x < W;
  const int yc = min(max(y, 0), H - 1), xc = min(max(x, 0), W - 1);
  bf16x8 v;
  if (c < C1) {
    v = *(const bf16x8*)(X1 + (((size_t)n * H + yc) * W + xc) * C1 + c);
  } else {
    const int H2 = H >> 1, W2 = W >> 1;
    v = *(const bf16x8*)(X2 + (((size_t)n * H2 + (yc >> 1)) * W2 + (xc >> 1)) * C2 + (c - C1));
  }
  return in ? v : zero8();
}

// ------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM over pixels x output channels, K = 9*Cin.
// Workgroup: 4 waves, output tile 4 rows x 64 pixels x NCO channels; wave w
// owns row w.  Per 32-channel chunk the input tile (6 x 66 pixels x 32
// channels, 25 KB) is staged in LDS; each 16-byte group of 8 channels of LDS
// pixel p sits at group slot g ^ ((p >> 1) & 3), which keeps the MFMA operand
// reads (16 consecutive pixels, one group) free of bank conflicts without
// padding.  Weights: 32-channel tiles read them straight from global memory
// (small and shared by every workgroup, so they stay in L1/L2), one tap ahead
// of their use; 64-channel tiles stage them in LDS one row of three taps at a
// time.  The weights are the MFMA A operand and the pixels the B operand, so
// a lane's accumulators hold four consecutive output channels of one pixel.
// Each wave issues 4 x (NCO/16) MFMAs per tap.
// MODE 0: y = relu(acc + bias).  MODE 1: y = acc * (mask > 0) (mask may be
// null: y = acc).  Used for the forward pass (weights [Cout][9][Cin]) and
// for the input gradient (dY as input, flipped transposed weights).
// LDS rows of the staged tile are kInPitch = 72 pixels apart (a multiple of
// 8), so the bank swizzle of a fragment depends only on the lane and dx: the
// fragment reads are immediate offsets from three per-lane bases, and the
// staging slots, their LDS offsets and (away from the image border) their
// global offsets are per-thread constants.
constexpr int kTileW = 64, kTileH = 4;
constexpr int kInW = kTileW + 2, kInPitch = kTileW + 8;
constexpr int kInSlots = (kTileH + 2) * kInPitch * 4;  // 16-byte slots of the staged tile

__device__ __forceinline__ int swz(int p, int g) { return p * 32 + ((g ^ ((p >> 1) & 3)) << 3); }
// byte offset of 16-byte group g of LDS pixel p (the same swizzle)
__device__ __forceinline__ int swz_b(int p, int g) { return p * 64 + ((g ^ ((p >> 1) & 3)) << 4); }

template <int NCO, int MODE>
__global__ __launch_bounds__(256) void conv3_mfma_kernel(
    const __bf16* __restrict__ X1, int C1, const __bf16* __restrict__ X2, int C2,
    const __bf16* __restrict__ Wt, const float* __restrict__ bias,
    const __bf16* __restrict__ mask, __bf16* __restrict__ Y, int n, int H, int W, int Cout) {
  constexpr int NB = NCO / 16;
  // 64-channel tiles read each weight fragment from 4 waves: stage the weights
  // in LDS, one row of 3 taps at a time; 32-channel tiles read them from
  // global memory (L1), which leaves LDS for more resident workgroups
  constexpr bool kWLds = NCO >= 64;
  __shared__ __attribute__((aligned(16))) char sIn[kInSlots * 16];
  __shared__ __attribute__((aligned(16))) char sW[kWLds ? 3 * NCO * 64 : 16];
  const int Cin = C1 + C2;
  const int tilesx = (W + kTileW - 1) / kTileW;
  const int tilesy = H / kTileH;
  const int tile = blockIdx.x;
  const int img = tile / (tilesx * tilesy);
  const int ty = (tile / tilesx) % tilesy, tx = tile % tilesx;
  const int y0 = ty * kTileH, x0 = tx * kTileW;
  const int co0 = blockIdx.y * NCO;
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int lr = lane % 16, lk = lane / 16;

  // staging slot r of this thread: LDS pixel sp = tid/4 + 64 r, group sg =
  // tid % 4, at byte sbase + 4096 r (the swizzle term of sp is (tid >> 3) & 3
  // for every r); live bit r: the slot is a tile pixel, not row padding
  constexpr int kLd = (kInSlots + 255) / 256;
  const int sg = tid & 3;
  const int sbase = swz_b(tid >> 2, sg);
  const bool interior = y0 >= 1 && y0 + kTileH + 1 <= H && x0 >= 1 && x0 + kTileW + 1 <= W;
  unsigned live = 0;
  int off1[kLd];  // X1 element offset from the tile's corner (interior tiles)
#pragma unroll
  for (int r = 0; r < kLd; ++r) {
    const int sp = (tid >> 2) + 64 * r, ry = sp / kInPitch, rx = sp % kInPitch;
    live |= (sp < kInSlots / 4 && rx < kInW ? 1u : 0u) << r;
    off1[r] = (ry * W + rx) * C1 + sg * 8;
  }
  // fragment bases: pixel (wave + dy, 16 mb + lr + dx) of the tile sits at
  // pbase[dx] + (dy * kInPitch + 16 mb) * 64
  int pbase[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) pbase[dx] = swz_b(wave * kInPitch + lr + dx, lk);

  floatx4 acc[4][NB];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  // weight operand of block nb at tap t of chunk ci0: channel co0 + 16 nb + lr,
  // input channels ci0 + 8 lk .. + 7
  auto load_b = [&](bf16x8* b, int t, int ci0) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = co0 + nb * 16 + lr;
      b[nb] = co < Cout ? *(const bf16x8*)(Wt + ((size_t)co * 9 + t) * Cin + ci0 + lk * 8)
                        : zero8();
    }
  };
  auto tap = [&](int t, const bf16x8* b) {
    const int dy = t / 3, dx = t % 3;  // offsets + 1
    bf16x8 a[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
      a[mb] = *(const bf16x8*)(sIn + pbase[dx] + (dy * kInPitch + mb * 16) * 64);
    // weights as the A operand, pixels as B: the 16 x 16 result is
    // [output channel][pixel], so each lane ends with 4 consecutive channels
    // of one pixel (one 8-byte NHWC store in the epilogue)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[nb], a[mb], acc[mb][nb], 0, 0, 0);
  };

  // weight staging (NCO >= 64): one row of three taps, 3 * NCO * 32 bf16, as
  // kWl 16-byte loads per thread, all issued before any is stored; the next
  // row's loads are in flight during the current row's MFMAs
  // weight slot r (tap tl of the row, channel co = sp / NCO, sp % NCO) at the
  // same LDS offsets as the tile's slots; its global offset less 3 tg Cin + ci0
  constexpr int kWl = (3 * NCO * 4 + 255) / 256;
  bf16x8 wl[kWLds ? kWl : 1];
  int offw[kWLds ? kWl : 1];
  unsigned wlive = 0;
  if constexpr (kWLds) {
#pragma unroll
    for (int r = 0; r < kWl; ++r) {
      const int sp = (tid >> 2) + 64 * r, tl = min(sp / NCO, 2), co = sp % NCO;
      wlive |= (sp < 3 * NCO && co0 + co < Cout ? 1u : 0u) << r;
      offw[r] = (min(co0 + co, Cout - 1) * 9 + tl) * Cin + sg * 8;
    }
  }
  const int wbase = swz_b(lr, lk);
  auto load_w = [&](int tg, int ci0) {
#pragma unroll
    for (int r = 0; r < kWl; ++r) {
      const bf16x8 v = *(const bf16x8*)(Wt + offw[r] + 3 * tg * Cin + ci0);
      wl[r] = (wlive >> r) & 1 ? v : zero8();
    }
  };
  auto store_w = [&]() {
#pragma unroll
    for (int r = 0; r < kWl; ++r)
      if ((tid >> 2) + 64 * r < 3 * NCO) *(bf16x8*)(sW + sbase + 4096 * r) = wl[r];
  };

  for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
    bf16x8 b[NB];
    if (!kWLds) load_b(b, 0, ci0);
    if constexpr (kWLds) load_w(0, ci0);
    // the chunk's input tile: every 16-byte load of this thread is issued
    // before the first is stored (one memory round trip, not one per load)
    bf16x8 ld[kLd];
    if (interior && ci0 < C1) {
      const __bf16* base = X1 + (((size_t)img * H + (y0 - 1)) * W + (x0 - 1)) * C1 + ci0;
#pragma unroll
      for (int r = 0; r < kLd; ++r) ld[r] = *(const bf16x8*)(base + off1[r]);
    } else {
#pragma unroll
      for (int r = 0; r < kLd; ++r) {
        const int sp = min((tid >> 2) + 64 * r, kInSlots / 4 - 1);
        const int ry = sp / kInPitch, rx = min(sp % kInPitch, kInW - 1);
        ld[r] = load8_clamped(X1, C1, X2, C2, img, y0 + ry - 1, x0 + rx - 1, H, W, ci0 + sg * 8);
      }
    }
    __syncthreads();  // previous chunk fully consumed
#pragma unroll
    for (int r = 0; r < kLd; ++r)
      if ((live >> r) & 1) *(bf16x8*)(sIn + sbase + 4096 * r) = ld[r];
    if constexpr (kWLds) {
#pragma unroll
      for (int tg = 0; tg < 3; ++tg) {
        if (tg > 0) __syncthreads();  // previous tap row consumed
        store_w();
        if (tg < 2) load_w(tg + 1, ci0);
        __syncthreads();
#pragma unroll
        for (int tl = 0; tl < 3; ++tl) {
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
            b[nb] = *(const bf16x8*)(sW + wbase + (tl * NCO + nb * 16) * 64);
          tap(3 * tg + tl, b);
        }
      }
    } else {
      __syncthreads();
#pragma unroll 1
      for (int t = 0; t < 9; ++t) {
        bf16x8 bn[NB];
        if (t < 8) load_b(bn, t + 1, ci0);
        tap(t, b);
        if (t < 8) {
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) b[nb] = bn[nb];
        }
      }
    }
  }
  // epilogue: lane holds rows 4*lk + j (output channels) of column lr (pixel)
  const int y = y0 + wave;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int cb = co0 + nb * 16 + 4 * lk;  // channels cb .. cb + 3 (Cout % 16 == 0)
    if (cb >= Cout) continue;
    float bv[4] = {0.f, 0.f, 0.f, 0.f};
    if (MODE == 0)
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = bias[cb + j];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const int x = x0 + mb * 16 + lr;
      if (x >= W) continue;
      const size_t o = (((size_t)img * H + y) * W + x) * Cout + cb;
      bf16x4 out;
      bf16x4 mk;
      if (MODE == 1 && mask) mk = *(const bf16x4*)(mask + o);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float v = acc[mb][nb][j];
        if (MODE == 0) {
          v = fmaxf(v + bv[j], 0.f);
        } else if (mask) {
          v = ((float)mk[j] > 0.f) ? v : 0.f;
        }
        out[j] = (__bf16)v;
      }
      *(bf16x4*)(Y + o) = out;
    }
  }
}

// ------------------------------------------------------------------------
// 3x3 convolution, the same implicit GEMM, persistent and software-pipelined,
// with its addressing reduced to per-thread constants (the kernel above spends
// more issue cycles on address arithmetic than on MFMAs).  Used for 32 output
// channels (the full-resolution layers); with 64-channel tiles the per-item
// restaging of 9 x 64 x 32 weights made it slower than the kernel above.
// Workgroup: 4 waves in a 2 x 2 grid, output tile 8 rows x 32 PX pixels x NCO
// channels (NCO = 32: PX = 2); wave (wr, wc) owns rows
// 4 wr .. 4 wr + 3 and pixels 16 PX wc .. + 16 PX - 1.  The workgroup runs the
// items (tile, 32-channel chunk) of tiles b, b + G, ... (G workgroups); the
// next item's input tile (10 x (32 PX + 2) pixels) and weights (9 x NCO x 32,
// loaded once when there is a single chunk) are read into registers while the
// current item's MFMAs run and stored to LDS after them, two barriers per
// item.  LDS rows are 32 PX + 8 pixels apart (a multiple of 8), so the bank
// swizzle of a fragment depends only on the lane and dx and every fragment
// read is an immediate offset from one of three per-lane bases; staging slots
// and their global offsets are per-thread constants, and tiles away from the
// image border load without clamping.  The taps run by column dx: each of the
// wave's 6 input rows is read once per dx (PX 16-pixel fragments) and applied
// to every output row that uses it (dy = input row - output row).
// Same modes and results (up to the FP32 summation order) as the kernel above.
constexpr int kRTileH = 8, kRInRows = kRTileH + 2;
template <int NCO>
struct RowsConv {
  static constexpr int PX = NCO == 32 ? 2 : 1;       // 16-pixel blocks per wave row
  static constexpr int Q = NCO / 16;                 // 16-channel blocks
  static constexpr int TW = 32 * PX;                 // tile width (pixels)
  static constexpr int InW = TW + 2;                 // staged pixels per row
  static constexpr int Pitch = TW + 8;               // LDS pixels per tile row
  static constexpr int Slots = kRInRows * Pitch * 4; // 16-byte LDS slots of the tile
  static constexpr int Ld = (Slots + 255) / 256;     // per thread
  static constexpr int WSlots = 9 * NCO * 4;         // weight slots (taps x co x groups)
  static constexpr int Wl = (WSlots + 255) / 256;
};

// 2 x 2 max-pool of a ReLU conv output still held as MFMA accumulators: rows
// o and o + 1 are two registers of the lane (a, b), pixels x and x + 1 two
// adjacent lanes (lr even / odd).  Every value is rounded to bf16 before the
// maximum, so the result equals maxpool2_kernel's on the stored output.  All
// 64 lanes must call it (lane exchange); the even lanes hold the result.
__device__ __forceinline__ bf16x4 pool_pair(const floatx4& a, const floatx4& b,
                                            const float (&bv)[4]) {
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float u = (float)(__bf16)fmaxf(a[j] + bv[j], 0.f);
    const float v = (float)(__bf16)fmaxf(b[j] + bv[j], 0.f);
    float m = fmaxf(u, v);
    m = fmaxf(m, __shfl_xor(m, 1));
    r[j] = (__bf16)m;
  }
  return r;
}

template <int NCO, int MODE>
__global__ __launch_bounds__(256, 2) void conv3_rows_kernel(
    const __bf16* __restrict__ X1, int C1, const __bf16* __restrict__ X2, int C2,
    const __bf16* __restrict__ Wt, const float* __restrict__ bias,
    const __bf16* __restrict__ mask, __bf16* __restrict__ Y, int n, int H, int W, int Cout,
    __bf16* __restrict__ pool) {
  static_assert(NCO == 32, "64-channel tiles use conv3_res64_kernel or conv3_mfma_kernel");
  using K = RowsConv<NCO>;
  constexpr int Q = K::Q, P = K::PX, R = 4;
  constexpr int kRTileW = K::TW, kRInW = K::InW, kRPitch = K::Pitch;
  constexpr int kRLd = K::Ld, kRWl = K::Wl;
  __shared__ __attribute__((aligned(16))) char sIn[kRInRows * kRPitch * 64];
  __shared__ __attribute__((aligned(16))) char sW[K::WSlots * 16];
  const int Cin = C1 + C2;
  const bool one_chunk = Cin == 32;
  const int tilesx = (W + kRTileW - 1) / kRTileW;
  const int tilesy = H / kRTileH;
  const int ntiles = n * tilesy * tilesx;
  const int co0 = blockIdx.y * NCO;
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int wr = wave / 2, wc = wave % 2;
  const int lr = lane % 16, lk = lane / 16;

  // staging slot r of this thread: LDS pixel sp = tid/4 + 64 r, group sg =
  // tid % 4; its LDS byte offset is sbase + 4096 r (the swizzle term of sp is
  // (tid >> 3) & 3 for every r)
  const int sg = tid & 3;
  const int sbase = swz_b(tid >> 2, sg);
  unsigned live = 0;   // bit r: slot r is a tile pixel (not row padding / past the end)
  int off1[kRLd];      // X1 element offset from the item's corner (interior tiles)
#pragma unroll
  for (int r = 0; r < kRLd; ++r) {
    const int sp = (tid >> 2) + 64 * r, ry = sp / kRPitch, rx = sp % kRPitch;
    const bool ok = sp < kRInRows * kRPitch && rx < kRInW;
    live |= (ok ? 1u : 0u) << r;
    off1[r] = (ry * W + rx) * C1 + sg * 8;
  }
  // weight slot r: (tap t, channel co) = sp / 32, sp % 32 at the same LDS offsets
  int offw[kRWl];
  unsigned wlive = 0;
#pragma unroll
  for (int r = 0; r < kRWl; ++r) {
    const int sp = (tid >> 2) + 64 * r, t = sp / NCO, co = sp % NCO;
    const bool ok = sp < 9 * NCO && co0 + co < Cout;
    wlive |= (ok ? 1u : 0u) << r;
    offw[r] = ((min(co0 + co, Cout - 1) * 9 + min(t, 8)) * Cin) + sg * 8;
  }

  bf16x8 ld[kRLd], wl[kRWl];
  auto load_w = [&](int ci0) {
#pragma unroll
    for (int r = 0; r < kRWl; ++r) {
      const bf16x8 v = *(const bf16x8*)(Wt + offw[r] + ci0);
      wl[r] = (wlive >> r) & 1 ? v : zero8();
    }
  };
  auto load_item = [&](int tile, int ci0, bool with_w) {
    const int img = tile / (tilesx * tilesy);
    const int y0 = ((tile / tilesx) % tilesy) * kRTileH, x0 = (tile % tilesx) * kRTileW;
    const bool interior = y0 >= 1 && y0 + kRTileH + 1 <= H && x0 >= 1 && x0 + kRTileW + 1 <= W;
    if (interior && ci0 < C1) {
      const __bf16* base = X1 + (((size_t)img * H + (y0 - 1)) * W + (x0 - 1)) * C1 + ci0;
#pragma unroll
      for (int r = 0; r < kRLd; ++r) ld[r] = *(const bf16x8*)(base + off1[r]);
    } else {
#pragma unroll
      for (int r = 0; r < kRLd; ++r) {
        const int sp = min((tid >> 2) + 64 * r, kRInRows * kRPitch - 1);
        const int ry = sp / kRPitch, rx = min(sp % kRPitch, kRInW - 1);
        ld[r] = load8_clamped(X1, C1, X2, C2, img, y0 + ry - 1, x0 + rx - 1, H, W, ci0 + sg * 8);
      }
    }
    if (with_w) load_w(ci0);
  };
  auto store_item = [&](bool with_w) {
#pragma unroll
    for (int r = 0; r < kRLd; ++r)
      if ((live >> r) & 1) *(bf16x8*)(sIn + sbase + 4096 * r) = ld[r];
    if (with_w) {
#pragma unroll
      for (int r = 0; r < kRWl; ++r)
        if ((tid >> 2) + 64 * r < 9 * NCO) *(bf16x8*)(sW + sbase + 4096 * r) = wl[r];
    }
  };

  // fragment bases: pixel (4 wr + r, 16 P wc + 16 p + lr + dx) of the tile sits at
  // pbase[dx] + (r * kRPitch + 16 p) * 64; weight (tap t, 16 q + lr) at
  // wbase + (t * NCO + 16 q) * 64
  int pbase[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
    pbase[dx] = swz_b(4 * wr * kRPitch + 16 * P * wc + lr + dx, lk);
  const int wbase = swz_b(lr, lk);

  int tile = blockIdx.x;  // uniform per workgroup: every barrier below is reached by all waves
  if (tile >= ntiles) return;
  load_item(tile, 0, true);
  bool first = true;
  for (; tile < ntiles; tile += gridDim.x) {
    floatx4 acc[R][P][Q];
#pragma unroll
    for (int o = 0; o < R; ++o)
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[o][p][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
      const bool w_now = first || !one_chunk;
      __syncthreads();  // the previous item's LDS reads are done
      store_item(w_now);
      __syncthreads();
      first = false;
      {  // the next item's loads, in flight during this item's MFMAs
        int nt = tile, nc = ci0 + 32;
        if (nc >= Cin) {
          nt = tile + gridDim.x;
          nc = 0;
        }
        if (nt < ntiles) load_item(nt, nc, !one_chunk);
      }
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        // weights as the A operand, pixels as B (lane: 4 channels of one pixel)
        bf16x8 wf[3][Q];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
          for (int q = 0; q < Q; ++q)
            wf[dy][q] = *(const bf16x8*)(sW + wbase + ((3 * dy + dx) * NCO + 16 * q) * 64);
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {  // input row 4 wr + r of the tile
          bf16x8 pf[P];
#pragma unroll
          for (int p = 0; p < P; ++p)
            pf[p] = *(const bf16x8*)(sIn + pbase[dx] + (r * kRPitch + 16 * p) * 64);
#pragma unroll
          for (int o = 0; o < R; ++o) {
            const int dy = r - o;
            if (dy < 0 || dy > 2) continue;
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
              for (int q = 0; q < Q; ++q)
                acc[o][p][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[dy][q], pf[p],
                                                                       acc[o][p][q], 0, 0, 0);
          }
        }
      }
    }
    // epilogue: lane holds channels 4 lk .. 4 lk + 3 of block q, pixel lr of block p
    const int img = tile / (tilesx * tilesy);
    const int y0 = ((tile / tilesx) % tilesy) * kRTileH, x0 = (tile % tilesx) * kRTileW;
    const bool full = x0 + kRTileW <= W && co0 + NCO <= Cout;
    const size_t tbase = (((size_t)img * H + y0 + 4 * wr) * W + x0 + 16 * P * wc) * Cout + co0;
    float bv[Q][4];  // read per tile: kept out of the registers of the MFMA loop
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bv[q][j] = MODE == 0 ? bias[min(co0 + q * 16 + 4 * lk + j, Cout - 1)] : 0.f;
#pragma unroll
    for (int o = 0; o < R; ++o) {
      __bf16* yrow = Y + tbase + (size_t)o * W * Cout;
      const __bf16* mrow = MODE == 1 && mask ? mask + tbase + (size_t)o * W * Cout : nullptr;
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int lo = (16 * p + lr) * Cout + 16 * q + 4 * lk;
          if (!full && (x0 + 16 * P * wc + 16 * p + lr >= W || co0 + 16 * q + 4 * lk >= Cout))
            continue;
          bf16x4 out;
          bf16x4 mk;
          if (MODE == 1 && mask) mk = *(const bf16x4*)(mrow + lo);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            float v = acc[o][p][q][j];
            if (MODE == 0) {
              v = fmaxf(v + bv[q][j], 0.f);
            } else if (mask) {
              v = ((float)mk[j] > 0.f) ? v : 0.f;
            }
            out[j] = (__bf16)v;
          }
          *(bf16x4*)(yrow + lo) = out;
        }
    }
    if constexpr (MODE == 0) {
      if (pool) {  // fused 2 x 2 max-pool (H, W even; rows y0 + 4 wr + o, o even, pair up)
        const int H2 = H >> 1, W2 = W >> 1;
#pragma unroll
        for (int o = 0; o < R; o += 2) {
          const int y2 = (y0 + 4 * wr + o) >> 1;
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const int x = x0 + 16 * P * wc + 16 * p + lr;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
              const int cb = co0 + 16 * q + 4 * lk;
              const bf16x4 m = pool_pair(acc[o][p][q], acc[o + 1][p][q], bv[q]);
              if (!(lr & 1) && x < W && cb < Cout)
                *(bf16x4*)(pool + (((size_t)img * H2 + y2) * W2 + (x >> 1)) * Cout + cb) = m;
            }
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// 3x3 convolution for 64 output channels with the workgroup's weights in LDS.
// Up to 64 input channels (STREAM = false) every weight of the workgroup's 64
// output channels stays resident for the whole kernel (9 x 64 x Cin bf16,
// loaded once), so that the per-tile work is the input tile only; above that
// (STREAM = true) LDS holds one 32-channel chunk of weights, loaded into
// registers together with the input tile of the next item and stored after the
// same barrier.  Workgroup: 8 waves, 4 x 2 grid, output tile 16 rows x 32
// pixels x 64 channels; wave (wr, wc) owns rows 4 wr .. 4 wr + 3 and pixels
// 16 wc .. 16 wc + 15.  Persistent over items (tile, 32-channel chunk) like
// conv3_rows_kernel, the next item's 18 x 34-pixel tile loaded into registers
// during the current item's MFMAs; LDS rows 40 pixels apart, so fragment reads
// are immediate offsets from per-lane bases.  One workgroup per CU (83-120 KB
// of LDS), two waves per SIMD.  Same modes and results (up to the FP32
// summation order) as conv3_mfma_kernel.
constexpr int kXTileH = 16, kXTileW = 32, kXInRows = kXTileH + 2, kXInW = kXTileW + 2,
              kXPitch = kXTileW + 8;
constexpr int kXSlots = kXInRows * kXPitch * 4;  // 16-byte slots of the tile
constexpr int kXLd = (kXSlots + 511) / 512;       // per thread
constexpr int kXMaxCin = 64;
constexpr int kXWLd = (9 * 64 * 4 + 511) / 512;   // weight slots of one chunk per thread

template <int MODE, bool STREAM>
__global__ __launch_bounds__(512, 1) void conv3_res64_kernel(
    const __bf16* __restrict__ X1, int C1, const __bf16* __restrict__ X2, int C2,
    const __bf16* __restrict__ Wt, const float* __restrict__ bias,
    const __bf16* __restrict__ mask, __bf16* __restrict__ Y, int n, int H, int W, int Cout,
    __bf16* __restrict__ pool) {
  constexpr int NCO = 64, Q = 4, R = 4;
  __shared__ __attribute__((aligned(16))) char sIn[kXInRows * kXPitch * 64];
  __shared__ __attribute__((aligned(16))) char sW[(STREAM ? 1 : kXMaxCin / 32) * 9 * NCO * 64];
  const int Cin = C1 + C2, nch = Cin / 32;
  const int tilesx = (W + kXTileW - 1) / kXTileW;
  const int tilesy = H / kXTileH;
  const int ntiles = n * tilesy * tilesx;
  const int co0 = blockIdx.y * NCO;
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int wr = wave / 2, wc = wave % 2;
  const int lr = lane % 16, lk = lane / 16;
  const int sg = tid & 3;
  // staging slot r of this thread: LDS pixel tid/4 + 128 r, group sg, at byte
  // sbase + 8192 r (the swizzle term of that pixel is (tid >> 3) & 3 for every r)
  const int sbase = swz_b(tid >> 2, sg);

  // all weights of the workgroup's channels: LDS pixel (c 9 + t) NCO + co
  if constexpr (!STREAM) {
    const int nslot = nch * 9 * NCO * 4;
    for (int i = tid; i < nslot; i += 512) {
      const int g = i & 3, px = i >> 2;
      const int co = px % NCO, ct = px / NCO, c = ct / 9, t = ct % 9;
      bf16x8 v = zero8();
      if (co0 + co < Cout) v = *(const bf16x8*)(Wt + ((size_t)(co0 + co) * 9 + t) * Cin + c * 32 + g * 8);
      *(bf16x8*)(sW + swz_b(px, g)) = v;
    }
  }

  unsigned live = 0;
  int off1[kXLd];
#pragma unroll
  for (int r = 0; r < kXLd; ++r) {
    const int sp = (tid >> 2) + 128 * r, ry = sp / kXPitch, rx = sp % kXPitch;
    live |= (sp < kXInRows * kXPitch && rx < kXInW ? 1u : 0u) << r;
    off1[r] = (ry * W + rx) * C1 + sg * 8;
  }
  // STREAM: weight slot r of this thread is LDS pixel tid/4 + 128 r of the
  // chunk, i.e. tap (tid >> 8) + 2 r of output channel (tid >> 2) & 63, group
  // sg, stored at the same sbase + 8192 r as the tile's slots
  const int wco = (tid >> 2) & 63;
  const bool wok = co0 + wco < Cout;
  int woff[kXWLd];
  unsigned wlive = 0;
#pragma unroll
  for (int r = 0; r < kXWLd; ++r) {
    const int t = (tid >> 8) + 2 * r;
    wlive |= (t < 9 && wok ? 1u : 0u) << r;
    woff[r] = ((co0 + wco) * 9 + min(t, 8)) * Cin + sg * 8;
  }
  bf16x8 ld[kXLd];
  bf16x8 wl[STREAM ? kXWLd : 1];
  auto load_item = [&](int tile, int ci0) {
    if constexpr (STREAM) {
#pragma unroll
      for (int r = 0; r < kXWLd; ++r)
        wl[r] = ((wlive >> r) & 1) ? *(const bf16x8*)(Wt + woff[r] + ci0) : zero8();
    }
    const int img = tile / (tilesx * tilesy);
    const int y0 = ((tile / tilesx) % tilesy) * kXTileH, x0 = (tile % tilesx) * kXTileW;
    const bool interior = y0 >= 1 && y0 + kXTileH + 1 <= H && x0 >= 1 && x0 + kXTileW + 1 <= W;
    if (interior && ci0 < C1) {
      const __bf16* base = X1 + (((size_t)img * H + (y0 - 1)) * W + (x0 - 1)) * C1 + ci0;
#pragma unroll
      for (int r = 0; r < kXLd; ++r) ld[r] = *(const bf16x8*)(base + off1[r]);
    } else {
      int t2 = tid;
      __asm__ volatile("" : "+v"(t2));
#pragma unroll
      for (int r = 0; r < kXLd; ++r) {
        const int sp = min((t2 >> 2) + 128 * r, kXInRows * kXPitch - 1);
        const int ry = sp / kXPitch, rx = min(sp % kXPitch, kXInW - 1);
        ld[r] = load8_clamped(X1, C1, X2, C2, img, y0 + ry - 1, x0 + rx - 1, H, W, ci0 + sg * 8);
      }
    }
  };

  int pbase[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) pbase[dx] = swz_b(4 * wr * kXPitch + 16 * wc + lr + dx, lk);
  const int wbase = swz_b(lr, lk);

  int tile = blockIdx.x;  // uniform per workgroup: every barrier below is reached by all waves
  if (tile >= ntiles) return;
  load_item(tile, 0);
  for (; tile < ntiles; tile += gridDim.x) {
    floatx4 acc[R][Q];
#pragma unroll
    for (int o = 0; o < R; ++o)
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[o][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nch; ++c) {
      __syncthreads();  // the previous item's LDS reads (and the weight stores) are done
#pragma unroll
      for (int r = 0; r < kXLd; ++r)
        if ((live >> r) & 1) *(bf16x8*)(sIn + sbase + 8192 * r) = ld[r];
      if constexpr (STREAM) {
#pragma unroll
        for (int r = 0; r < kXWLd; ++r)
          if ((tid >> 8) + 2 * r < 9) *(bf16x8*)(sW + sbase + 8192 * r) = wl[r];
      }
      __syncthreads();
      {  // the next item's loads, in flight during this item's MFMAs
        int nt = tile, nc = c + 1;
        if (nc >= nch) {
          nt = tile + gridDim.x;
          nc = 0;
        }
        if (nt < ntiles) load_item(nt, nc * 32);
      }
      const char* wch = STREAM ? sW : sW + c * 9 * NCO * 64;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        bf16x8 wf[3][Q];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
          for (int q = 0; q < Q; ++q)
            wf[dy][q] = *(const bf16x8*)(wch + wbase + ((3 * dy + dx) * NCO + 16 * q) * 64);
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
          const bf16x8 pf = *(const bf16x8*)(sIn + pbase[dx] + r * kXPitch * 64);
#pragma unroll
          for (int o = 0; o < R; ++o) {
            const int dy = r - o;
            if (dy < 0 || dy > 2) continue;
#pragma unroll
            for (int q = 0; q < Q; ++q)
              acc[o][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[dy][q], pf, acc[o][q], 0, 0, 0);
          }
        }
      }
    }
    const int img = tile / (tilesx * tilesy);
    const int y0 = ((tile / tilesx) % tilesy) * kXTileH, x0 = (tile % tilesx) * kXTileW;
    const int x = x0 + 16 * wc + lr;
    float bv[Q][4];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bv[q][j] = MODE == 0 ? bias[min(co0 + q * 16 + 4 * lk + j, Cout - 1)] : 0.f;
#pragma unroll
    for (int o = 0; o < R; ++o) {
      const int y = y0 + 4 * wr + o;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int cb = co0 + q * 16 + 4 * lk;
        if (x >= W || cb >= Cout) continue;
        const size_t off = (((size_t)img * H + y) * W + x) * Cout + cb;
        bf16x4 out;
        bf16x4 mk;
        if (MODE == 1 && mask) mk = *(const bf16x4*)(mask + off);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = acc[o][q][j];
          if (MODE == 0) {
            v = fmaxf(v + bv[q][j], 0.f);
          } else if (mask) {
            v = ((float)mk[j] > 0.f) ? v : 0.f;
          }
          out[j] = (__bf16)v;
        }
        *(bf16x4*)(Y + off) = out;
      }
    }
    if constexpr (MODE == 0) {
      if (pool) {  // fused 2 x 2 max-pool (H, W even; rows y0 + 4 wr + o, o even, pair up)
        const int H2 = H >> 1, W2 = W >> 1;
#pragma unroll
        for (int o = 0; o < R; o += 2) {
          const int y2 = (y0 + 4 * wr + o) >> 1;
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const int cb = co0 + q * 16 + 4 * lk;
            const bf16x4 m = pool_pair(acc[o][q], acc[o + 1][q], bv[q]);
            if (!(lr & 1) && x < W && cb < Cout)
              *(bf16x4*)(pool + (((size_t)img * H2 + y2) * W2 + (x >> 1)) * Cout + cb) = m;
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution: dW[co][t][ci] += sum_p dY[p][co] *
// X[p + off_t][ci], one tap per workgroup (blockIdx.z), a 64 x 64 (co, ci)
// tile (blockIdx.y) and a range of pixels (blockIdx.x, split K; partial sums
// are added to the FP32 gradient atomically).  32-pixel chunks of dY and of
// the shifted input are staged transposed in LDS ([channel][pixel]) so that
// both MFMA operands are 16-byte LDS reads of 8 consecutive pixels.
constexpr int kDwPix = 32, kDwPad = 40;

__global__ __launch_bounds__(256) void conv3_dw_kernel(
    const __bf16* __restrict__ dY, int Cout, const __bf16* __restrict__ X1, int C1,
    const __bf16* __restrict__ X2, int C2, float* __restrict__ dW, int n, int H, int W,
    int pix_per_block) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[64 * kDwPad];  // [co][pixel]
  __shared__ __attribute__((aligned(16))) __bf16 sB[64 * kDwPad];  // [ci][pixel]
  const int Cin = C1 + C2;
  const int t = blockIdx.z, dy = t / 3 - 1, dx = t % 3 - 1;
  const int nci = (Cin + 63) / 64;
  const int co0 = (blockIdx.y / nci) * 64, ci0 = (blockIdx.y % nci) * 64;
  const long P = (long)n * H * W;
  const long p0 = (long)blockIdx.x * pix_per_block;
  const long p1 = std::min(P, p0 + pix_per_block);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int lr = lane % 16, lk = lane / 16;
  const int wm = wave / 2, wn = wave % 2;  // wave tile: co [32 wm, +32), ci [32 wn, +32)
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // staging role of this thread: pixel sp of the chunk, 8-channel group sg
  const int sp = threadIdx.x / 8, sg = threadIdx.x % 8;
  for (long pc = p0; pc < p1; pc += kDwPix) {
    const long p = pc + sp;
    bf16x8 va = zero8(), vb = zero8();
    if (p < p1) {
      const int x = (int)(p % W), y = (int)((p / W) % H), img = (int)(p / ((long)W * H));
      const int co = co0 + sg * 8, ci = ci0 + sg * 8;
      if (co < Cout) va = *(const bf16x8*)(dY + (size_t)p * Cout + co);
      if (ci < Cin) vb = load8(X1, C1, X2, C2, img, y + dy, x + dx, H, W, ci);
    }
    __syncthreads();  // previous chunk consumed
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      sA[(sg * 8 + i) * kDwPad + sp] = va[i];
      sB[(sg * 8 + i) * kDwPad + sp] = vb[i];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[i] = *(const bf16x8*)(sA + (wm * 32 + i * 16 + lr) * kDwPad + lk * 8);
      b[i] = *(const bf16x8*)(sB + (wn * 32 + i * 16 + lr) * kDwPad + lk * 8);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  // acc[i][j][r]: co = co0 + 32 wm + 16 i + 4 lk + r, ci = ci0 + 32 wn + 16 j + lr
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = co0 + 32 * wm + 16 * i + 4 * lk + r;
        const int ci = ci0 + 32 * wn + 16 * j + lr;
        if (co < Cout && ci < Cin) atomicAdd(dW + ((size_t)co * 9 + t) * Cin + ci, acc[i][j][r]);
      }
}

// ------------------------------------------------------------------------
// Input layer: 2 channels (x = (cos, sin), float2) -> C0 channels, 3x3, relu.
// One thread per pixel computing all C0 outputs (the 9 neighbours are loaded
// once); weights [C0][9][2] and biases in LDS, read as broadcasts.
__global__ __launch_bounds__(256) void conv3_in_kernel(const float2* __restrict__ x,
                                                       const float* __restrict__ w,
                                                       const float* __restrict__ b,
                                                       __bf16* __restrict__ Y, int n, int H,
                                                       int W, int C0) {
  // the weights and biases are read at wave-uniform addresses straight from
  // global memory: the compiler turns them into scalar loads (SGPR operands of
  // the FMAs) instead of one LDS read per multiply-add
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long P = (long)n * H * W;
  if (p >= P) return;
  const int xx = (int)(p % W), yy = (int)((p / W) % H), img = (int)(p / ((long)W * H));
  // the nine taps are loaded from clamped (in-image) pixels without branches,
  // so all nine loads are in flight together, then zeroed outside the image
  float in[18];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int y2 = yy + t / 3 - 1, x2 = xx + t % 3 - 1;
    const bool inside = y2 >= 0 && y2 < H && x2 >= 0 && x2 < W;
    const int yc = min(max(y2, 0), H - 1), xc = min(max(x2, 0), W - 1);
    const float2 v = x[((size_t)img * H + yc) * W + xc];
    in[2 * t] = inside ? v.x : 0.f;
    in[2 * t + 1] = inside ? v.y : 0.f;
  }
  for (int g = 0; g < C0 / 8; ++g) {
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int co = g * 8 + k;
      float s = b[co];
#pragma unroll
      for (int i = 0; i < 18; ++i) s = fmaf(w[co * 18 + i], in[i], s);
      o[k] = (__bf16)fmaxf(s, 0.f);
    }
    *(bf16x8*)(Y + (size_t)p * C0 + g * 8) = o;
  }
}

// weight/bias gradient of the input layer: dW[co][t][c] = sum_p dY[p][co] x[p+off_t][c]
__global__ __launch_bounds__(1024) void conv3_in_dw_kernel(const float2* __restrict__ x,
                                                          const __bf16* __restrict__ dY,
                                                          float* __restrict__ dW,
                                                          float* __restrict__ db, int n, int H,
                                                          int W, int C0, int pix_per_block) {
  // thread k < C0*19 (k over blockIdx.y, threadIdx.x) owns one weight (or
  // bias) and sums it over the block's pixels
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  const long P = (long)n * H * W;
  const long p0 = (long)blockIdx.x * pix_per_block, p1 = std::min(P, p0 + pix_per_block);
  if (k >= C0 * 19) return;
  const bool isb = k >= C0 * 18;
  const int co = isb ? k - C0 * 18 : k / 18;
  const int t = (k % 18) / 2, c = k % 2;
  float s = 0.f;
  for (long p = p0; p < p1; ++p) {
    const float g = (float)dY[(size_t)p * C0 + co];
    if (isb) {
      s += g;
      continue;
    }
    const int xx = (int)(p % W), yy = (int)((p / W) % H), img = (int)(p / ((long)W * H));
    const int y2 = yy + t / 3 - 1, x2 = xx + t % 3 - 1;
    if (y2 >= 0 && y2 < H && x2 >= 0 && x2 < W) {
      const float2 v = x[((size_t)img * H + y2) * W + x2];
      s = fmaf(g, c ? v.y : v.x, s);
    }
  }
  atomicAdd(isb ? db + co : dW + k, s);
}

// ------------------------------------------------------------------------
// Output layer: out = x + W d + b (1x1, C0 -> 2), FP32.
__global__ __launch_bounds__(256) void conv1_out_kernel(const __bf16* __restrict__ d,
                                                        const float2* __restrict__ x,
                                                        const float* __restrict__ w,
                                                        const float* __restrict__ b,
                                                        float2* __restrict__ out, long P, int C0) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s0 = b[0], s1 = b[1];
  for (int c = 0; c < C0; c += 8) {
    const bf16x8 v = *(const bf16x8*)(d + (size_t)p * C0 + c);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float f = (float)v[k];
      s0 = fmaf(w[c + k], f, s0);
      s1 = fmaf(w[C0 + c + k], f, s1);
    }
  }
  const float2 xi = x[p];
  out[p] = make_float2(xi.x + s0, xi.y + s1);
}

// Loss L = mean |out - u|^2 and its backward through the output layer:
// g = 2 (out - u) / P; dd = (W^T g) * (d > 0); dW += g d^T; db += g.
// Per-block partial sums of the loss and of dW/db go out through atomics.
__global__ __launch_bounds__(256) void conv1_out_bwd_kernel(
    const __bf16* __restrict__ d, const float2* __restrict__ out, const float2* __restrict__ u,
    const float* __restrict__ w, __bf16* __restrict__ dd, float* __restrict__ dW,
    float* __restrict__ db, float* __restrict__ loss, long P, int C0) {
  extern __shared__ float red[];  // (2*C0 + 3) x 256 partials would be large: reduce per warp
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const float invP = 1.f / (float)P;
  float g0 = 0.f, g1 = 0.f, l = 0.f;
  if (p < P) {
    const float2 o = out[p], t = u[p];
    const float r0 = o.x - t.x, r1 = o.y - t.y;
    l = (r0 * r0 + r1 * r1) * invP;
    g0 = 2.f * r0 * invP;
    g1 = 2.f * r1 * invP;
  }
  // loss and bias gradients: block reduction
  float* s = red;
  s[threadIdx.x] = l;
  s[256 + threadIdx.x] = g0;
  s[512 + threadIdx.x] = g1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[threadIdx.x] += s[threadIdx.x + o];
      s[256 + threadIdx.x] += s[256 + threadIdx.x + o];
      s[512 + threadIdx.x] += s[512 + threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(loss, s[0]);
    atomicAdd(db, s[256]);
    atomicAdd(db + 1, s[512]);
  }
  __syncthreads();
  // per channel: dd and the dW partial (block-reduced through LDS)
  for (int c = 0; c < C0; ++c) {
    float f = 0.f;
    if (p < P) {
      f = (float)d[(size_t)p * C0 + c];
      const float gd = w[c] * g0 + w[C0 + c] * g1;
      dd[(size_t)p * C0 + c] = (__bf16)(f > 0.f ? gd : 0.f);
    }
    s[threadIdx.x] = g0 * f;
    s[256 + threadIdx.x] = g1 * f;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        s[threadIdx.x] += s[threadIdx.x + o];
        s[256 + threadIdx.x] += s[256 + threadIdx.x + o];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      atomicAdd(dW + c, s[0]);
      atomicAdd(dW + C0 + c, s[256]);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void maxpool2_kernel(const __bf16* __restrict__ x,
                                                       __bf16* __restrict__ y, int n, int H,
                                                       int W, int C) {
  const int H2 = H / 2, W2 = W / 2, G = C / 8;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)n * H2 * W2 * G) return;
  const int g = (int)(gid % G);
  const long q = gid / G;
  const int x2 = (int)(q % W2), y2 = (int)((q / W2) % H2), img = (int)(q / ((long)W2 * H2));
  const __bf16* b = x + (((size_t)img * H + 2 * y2) * W + 2 * x2) * C + g * 8;
  const bf16x8 v0 = *(const bf16x8*)b, v1 = *(const bf16x8*)(b + C);
  const bf16x8 v2 = *(const bf16x8*)(b + (size_t)W * C), v3 = *(const bf16x8*)(b + (size_t)W * C + C);
  bf16x8 o;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    o[k] = (__bf16)fmaxf(fmaxf((float)v0[k], (float)v1[k]), fmaxf((float)v2[k], (float)v3[k]));
  *(bf16x8*)(y + (size_t)q * C + g * 8) = o;
}

// Gradient into an encoder output e (C channels, H x W) from its two
// consumers, then through its relu:  g_e = (pool adjoint of gp (H/2 x W/2)
// routed to the first maximum of each window, row-major) + gskip (channels
// [0, C) of the decoder's concatenated input gradient, pitch Cs; may be
// null), times (e > 0).
__global__ __launch_bounds__(256) void pool_bwd_kernel(const __bf16* __restrict__ e,
                                                       const __bf16* __restrict__ gp,
                                                       const __bf16* __restrict__ gskip, int Cs,
                                                       __bf16* __restrict__ ge, int n, int H,
                                                       int W, int C) {
  const int H2 = H / 2, W2 = W / 2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)n * H2 * W2 * C) return;
  const int c = (int)(gid % C);
  const long q = gid / C;
  const int x2 = (int)(q % W2), y2 = (int)((q / W2) % H2), img = (int)(q / ((long)W2 * H2));
  const float g = gp ? (float)gp[q * C + c] : 0.f;
  float best = -1.f;
  int arg = 0;
  size_t idx[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int yy = 2 * y2 + k / 2, xx = 2 * x2 + k % 2;
    const size_t pix = ((size_t)img * H + yy) * W + xx;
    idx[k] = pix;
    const float v = (float)e[pix * C + c];
    if (k == 0 || v > best) {
      best = v;
      arg = k;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t pix = idx[k];
    float v = (k == arg) ? g : 0.f;
    if (gskip) v += (float)gskip[pix * Cs + c];
    ge[pix * C + c] = (__bf16)(((float)e[pix * C + c] > 0.f) ? v : 0.f);
  }
}

// Adjoint of the nearest x2 upsampling of a decoder input (channels [C1,
// C1+C) of the concatenated gradient gcat, full resolution, pitch C1 + C),
// times (a > 0) for the low-resolution activation a (C channels).
__global__ __launch_bounds__(256) void up_bwd_kernel(const __bf16* __restrict__ gcat, int C1,
                                                     const __bf16* __restrict__ a,
                                                     __bf16* __restrict__ ga, int n, int H2,
                                                     int W2, int C) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)n * H2 * W2 * C) return;
  const int c = (int)(gid % C);
  const long q = gid / C;
  const int x2 = (int)(q % W2), y2 = (int)((q / W2) % H2), img = (int)(q / ((long)W2 * H2));
  const int H = 2 * H2, W = 2 * W2, Cc = C1 + C;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t pix = ((size_t)img * H + 2 * y2 + k / 2) * W + 2 * x2 + k % 2;
    s += (float)gcat[pix * Cc + C1 + c];
  }
  ga[q * C + c] = (__bf16)(((float)a[q * C + c] > 0.f) ? s : 0.f);
}

// db[c] += sum_p g[p][c]
__global__ __launch_bounds__(256) void bias_grad_kernel(const __bf16* __restrict__ g,
                                                        float* __restrict__ db, long P, int C,
                                                        int pix_per_block) {
  const int c = threadIdx.x;
  if (c >= C) return;
  const long p0 = (long)blockIdx.x * pix_per_block, p1 = std::min(P, p0 + pix_per_block);
  float s = 0.f;
  for (long p = p0; p < p1; ++p) s += (float)g[(size_t)p * C + c];
  atomicAdd(db + c, s);
}

// Adam on FP32 master weights; refreshes the BF16 copies: Wf [Cout][9][Cin]
// and, for 3x3 layers, the flipped transpose Wt [Cin][9][Cout].
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   long count, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ __launch_bounds__(256) void weights_to_bf16_kernel(const float* __restrict__ w,
                                                              __bf16* __restrict__ wf,
                                                              __bf16* __restrict__ wt, int Cout,
                                                              int Cin) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)Cout * 9 * Cin) return;
  const int ci = (int)(i % Cin), t = (int)((i / Cin) % 9), co = (int)(i / (9L * Cin));
  const __bf16 h = (__bf16)w[i];
  wf[i] = h;
  if (wt) wt[((size_t)ci * 9 + (8 - t)) * Cout + co] = h;
}

__global__ void fill_kernel(float* __restrict__ p, long count, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// ======================================================== host launchers
static unsigned blocks(long work, int per = 256) { return (unsigned)((work + per - 1) / per); }

hipError_t unet_conv3(const __bf16* X1, int C1, const __bf16* X2, int C2, const __bf16* Wt,
                      const float* bias, const __bf16* mask, __bf16* Y, int n, int H, int W,
                      int Cout, int mode, hipStream_t s, __bf16* pool) {
  if ((C1 % 32) || (C2 % 32) || (Cout % 16) || (H % kTileH) || (C2 && ((H | W) & 1)) ||
      (pool && (mode != 0 || ((H | W) & 1))))
    return hipErrorInvalidValue;
  const int nco = Cout >= 64 ? 64 : 32;
  const int cob = (Cout + nco - 1) / nco;
  if (nco == 32 && H % kRTileH == 0) {
    // persistent 8-row kernel: two workgroups per CU (64 KB LDS, <= 256 registers)
    static int cus = 0;
    if (cus == 0) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorInvalidValue;
    }
    const int rtiles = n * (H / kRTileH) * ((W + RowsConv<32>::TW - 1) / RowsConv<32>::TW);
    dim3 grid(std::min(rtiles, std::max(1, 2 * cus / cob)), cob);
    if (mode == 0)
      hipLaunchKernelGGL((conv3_rows_kernel<32, 0>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout, pool);
    else
      hipLaunchKernelGGL((conv3_rows_kernel<32, 1>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout, nullptr);
    return hipGetLastError();
  }
  if (nco == 64 && H % kXTileH == 0) {
    // one 8-wave workgroup per CU: weights resident up to 64 input channels;
    // above that, for the forward pass, streamed per chunk where the tiles fill
    // every CU
    static int cus = 0;
    if (cus == 0) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorInvalidValue;
    }
    const int xt = n * (H / kXTileH) * ((W + kXTileW - 1) / kXTileW);
    const bool resident = C1 + C2 <= kXMaxCin;
    if (resident || (mode == 0 && xt * cob >= cus)) {
      dim3 gridx(std::min(xt, std::max(1, cus / cob)), cob);
#define GLSU_CONV3_RES64(MD, ST)                                                                 \
  hipLaunchKernelGGL((conv3_res64_kernel<MD, ST>), gridx, dim3(512), 0, s, X1, C1, X2, C2, Wt,   \
                     bias, mask, Y, n, H, W, Cout, pool)
      if (resident) {
        if (mode == 0) GLSU_CONV3_RES64(0, false);
        else GLSU_CONV3_RES64(1, false);
      } else {
        GLSU_CONV3_RES64(0, true);
      }
#undef GLSU_CONV3_RES64
      return hipGetLastError();
    }
  }

  const int tiles = n * (H / kTileH) * ((W + kTileW - 1) / kTileW);
  dim3 grid(tiles, cob);
  if (nco == 64) {
    if (mode == 0)
      hipLaunchKernelGGL((conv3_mfma_kernel<64, 0>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout);
    else
      hipLaunchKernelGGL((conv3_mfma_kernel<64, 1>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((conv3_mfma_kernel<32, 0>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout);
    else
      hipLaunchKernelGGL((conv3_mfma_kernel<32, 1>), grid, dim3(256), 0, s, X1, C1, X2, C2, Wt,
                         bias, mask, Y, n, H, W, Cout);
  }
  // this kernel has no pooled epilogue
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess || !pool) return e;
  return unet_maxpool2(Y, pool, n, H, W, Cout, s);
}

hipError_t unet_conv3_dw(const __bf16* dY, int Cout, const __bf16* X1, int C1, const __bf16* X2,
                         int C2, float* dW, int n, int H, int W, hipStream_t s) {
  if ((C1 % 32) || (C2 % 32) || (Cout % 32)) return hipErrorInvalidValue;
  const long P = (long)n * H * W;
  const int ppb = 4096;
  const int tilesc = ((Cout + 63) / 64) * ((C1 + C2 + 63) / 64);
  dim3 grid((unsigned)((P + ppb - 1) / ppb), tilesc, 9);
  hipLaunchKernelGGL(conv3_dw_kernel, grid, dim3(256), 0, s, dY, Cout, X1, C1, X2, C2, dW, n, H,
                     W, ppb);
  return hipGetLastError();
}

hipError_t unet_conv3_in(const float2* x, const float* w, const float* b, __bf16* Y, int n,
                         int H, int W, int C0, hipStream_t s) {
  const long work = (long)n * H * W;
  hipLaunchKernelGGL(conv3_in_kernel, dim3(blocks(work)), dim3(256), 0, s, x, w, b, Y, n, H, W,
                     C0);
  return hipGetLastError();
}

hipError_t unet_conv3_in_dw(const float2* x, const __bf16* dY, float* dW, float* db, int n,
                            int H, int W, int C0, hipStream_t s) {
  const long P = (long)n * H * W;
  const int ppb = 2048;
  const int nk = C0 * 19;  // weights and biases
  const int threads = std::min(1024, ((nk + 63) / 64) * 64);
  hipLaunchKernelGGL(conv3_in_dw_kernel, dim3(blocks(P, ppb), (nk + threads - 1) / threads),
                     dim3(threads), 0, s, x, dY, dW, db, n, H, W, C0, ppb);
  return hipGetLastError();
}

hipError_t unet_conv1_out(const __bf16* d, const float2* x, const float* w, const float* b,
                          float2* out, long P, int C0, hipStream_t s) {
  hipLaunchKernelGGL(conv1_out_kernel, dim3(blocks(P)), dim3(256), 0, s, d, x, w, b, out, P, C0);
  return hipGetLastError();
}

hipError_t unet_conv1_out_bwd(const __bf16* d, const float2* out, const float2* u,
                              const float* w, __bf16* dd, float* dW, float* db, float* loss,
                              long P, int C0, hipStream_t s) {
  hipLaunchKernelGGL(conv1_out_bwd_kernel, dim3(blocks(P)), dim3(256), 3 * 256 * sizeof(float),
                     s, d, out, u, w, dd, dW, db, loss, P, C0);
  return hipGetLastError();
}

hipError_t unet_maxpool2(const __bf16* x, __bf16* y, int n, int H, int W, int C, hipStream_t s) {
  const long work = (long)n * (H / 2) * (W / 2) * (C / 8);
  hipLaunchKernelGGL(maxpool2_kernel, dim3(blocks(work)), dim3(256), 0, s, x, y, n, H, W, C);
  return hipGetLastError();
}

hipError_t unet_pool_bwd(const __bf16* e, const __bf16* gp, const __bf16* gskip, int Cs,
                         __bf16* ge, int n, int H, int W, int C, hipStream_t s) {
  const long work = (long)n * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(pool_bwd_kernel, dim3(blocks(work)), dim3(256), 0, s, e, gp, gskip, Cs, ge,
                     n, H, W, C);
  return hipGetLastError();
}

hipError_t unet_up_bwd(const __bf16* gcat, int C1, const __bf16* a, __bf16* ga, int n, int H2,
                       int W2, int C, hipStream_t s) {
  const long work = (long)n * H2 * W2 * C;
  hipLaunchKernelGGL(up_bwd_kernel, dim3(blocks(work)), dim3(256), 0, s, gcat, C1, a, ga, n, H2,
                     W2, C);
  return hipGetLastError();
}

hipError_t unet_bias_grad(const __bf16* g, float* db, long P, int C, hipStream_t s) {
  const int ppb = 1024;
  hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks(P, ppb)), dim3(((C + 63) / 64) * 64), 0, s, g,
                     db, P, C, ppb);
  return hipGetLastError();
}

hipError_t unet_adam(float* w, const float* g, float* m, float* v, long count, float lr,
                     float b1, float b2, float eps, int step, hipStream_t s) {
  const float c1 = 1.f / (1.f - powf(b1, (float)step));
  const float c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(count)), dim3(256), 0, s, w, g, m, v, count, lr,
                     b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t unet_weights_bf16(const float* w, __bf16* wf, __bf16* wt, int Cout, int Cin,
                             hipStream_t s) {
  hipLaunchKernelGGL(weights_to_bf16_kernel, dim3(blocks((long)Cout * 9 * Cin)), dim3(256), 0, s,
                     w, wf, wt, Cout, Cin);
  return hipGetLastError();
}

hipError_t unet_fill(float* p, long count, float v, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(blocks(count)), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

}  // namespace glsu

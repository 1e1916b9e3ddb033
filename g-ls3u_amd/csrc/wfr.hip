// wfr.hip -- WFR2 (windowed Fourier ridges), the WFR mode of the WFT2 classes
// (WFT/WFT2_CPU.cpp:391-694, WFT/WFT2_CUDAf.cu:1174-1319, MatLab/fp_wft2f.m:
// 103-147), on the separable WFF2 machinery.
//
// For every pair p = (wx, wy): sf_p = f * w_p (the first convolution of WFF2);
// per pixel the pair with the largest |sf_p| gives r, wx, wy and phase =
// angle(sf_p), the first such pair in (wy, wx) order on exact ties.  Then the
// curvatures cxx = -conv(wx_rep, x g)/sum(x^2 g), cyy = -conv(wy_rep, y g)/
// sum(y^2 g) of the border-replicated ridge maps, the compensated phase and
// the amplitude b.
//
//   rows      A_wx = f (*)_x hx_wx                  (the WFF2 row kernel)
//   columns   for each (wx, column): FFT_y once, then for every wy one
//             product and IFFT_y, keeping per pixel the first wy with the
//             largest |sf|^2 in registers; writes that sf and its wy index
//             (D planes and a u16 plane, tiled like A)
//   ridge     per pixel, the best over wx (ties to the smaller wy, then wx)
//   curvature g = gx gy is separable and the reference's FFT evaluation never
//             wraps at the sampled points, so each curvature is a direct
//             row pass then column pass with clamped (replicated) borders;
//             the compensation is fused into the column pass.
#include "wff_kernels.h"

namespace glsu {

__device__ __forceinline__ float mag2(float2 v) { return fmaf(v.x, v.x, v.y * v.y); }

constexpr uint16_t kNoPair = 0xFFFF;

template <class P>
__device__ __forceinline__ void cols_wfr_body(const float2* __restrict__ A,
                                              float2* __restrict__ D, uint16_t* __restrict__ I,
                                              const float* __restrict__ Hy,
                                              const float2* __restrict__ tw_g, int H, int nwy,
                                              size_t plane, int k, int x, int lane,
                                              const float2* ttab, float2* buf) {
  constexpr int T = P::T, E = P::E, N = P::N;
  using Ctx = FftLane<P, true>;
  Ctx ctx;
  ctx.init(lane, tw_g, ttab);
  const float2* colA = A + plane * k;
  float2 ah[E], best[E];
  int bidx[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int y = lane + T * j;
    ah[j] = y < H ? colA[tile_index(y, x, H)] : make_float2(0.f, 0.f);
    best[j] = make_float2(0.f, 0.f);
    bidx[j] = kNoPair;
  }
  fft_fwd<P>(ah, buf, ctx);
  for (int q = 0; q < nwy; ++q) {
    const float* hy = Hy + (size_t)q * N + lane;  // lane-major window spectra
    float2 v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) v[j] = rmul(ah[j], hy[j * T]);
    fft_inv<P>(v, buf, ctx);
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (mag2(v[j]) > mag2(best[j])) {  // strict: the first wy keeps a tie
        best[j] = v[j];
        bidx[j] = q;
      }
    }
  }
  float2* colD = D + plane * k;
  uint16_t* colI = I + plane * k;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int y = lane + T * j;
    if (y < H) {
      colD[tile_index(y, x, H)] = best[j];
      colI[tile_index(y, x, H)] = (uint16_t)bidx[j];
    }
  }
}

template <class P, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void wfr_cols_kernel(
    const float2* __restrict__ A, float2* __restrict__ D, uint16_t* __restrict__ I,
    const float* __restrict__ Hy, const float2* __restrict__ tw_g, int W, int H, int nwx,
    int nwy, size_t plane) {
  constexpr int T = P::T, E = P::E;
  extern __shared__ float2 smem[];
  float2* ttab = smem;
  FftLane<P, true>::fill_ttab(ttab, tw_g);
  __syncthreads();
  const int slot = threadIdx.x / T;
  float2* buf = smem + E * T + slot * P::LDS;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int x = (b / nwx) * (blockDim.x / T) + slot;
  if (x >= W) return;
  cols_wfr_body<P>(A, D, I, Hy, tw_g, H, nwy, plane, b % nwx, x, threadIdx.x % T, ttab, buf);
}

// Per pixel, in the tiled order of the planes: the best pair over wx.
__global__ __launch_bounds__(256) void wfr_ridge_kernel(
    const float2* __restrict__ D, const uint16_t* __restrict__ I, int W, int H, int nwx,
    size_t plane, double wxl, double wxi, double wyl, double wyi, WfrOut z) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  const size_t per_tile = (size_t)kTileCols * H;
  const int xb = (int)(t / per_tile), rem = (int)(t % per_tile);
  const int y = rem / kTileCols, x = xb * kTileCols + rem % kTileCols;
  if (x >= W) return;
  float bm = 0.f;
  int bk = -1, bq = kNoPair;
  float2 bs = make_float2(0.f, 0.f);
  for (int k = 0; k < nwx; ++k) {
    const int q = I[plane * k + t];
    if (q == kNoPair) continue;
    const float2 s = D[plane * k + t];
    const float m = mag2(s);
    if (m > bm || (m == bm && bk >= 0 && q < bq)) {
      bm = m;
      bk = k;
      bq = q;
      bs = s;
    }
  }
  const size_t id = (size_t)y * W + x;
  if (bk >= 0) {
    z.r[id] = sqrtf(bm);
    z.wx[id] = (float)(wxl + (double)bk * wxi);
    z.wy[id] = (float)(wyl + (double)bq * wyi);
    z.phase[id] = atan2f(bs.y, bs.x);
  } else {  // |sf| = 0 for every pair: nothing updates the zero start
    z.r[id] = 0.f;
    z.wx[id] = 0.f;
    z.wy[id] = 0.f;
    z.phase[id] = 0.f;
  }
}

// Row pass of the curvature fits: R(i, j) = (sum_k k gx(k) wx(i, j-k),
// sum_k gx(k) wy(i, j-k)), column index clamped into the image.
// Grid (ceil(W/256), H); LDS: both rows' segments with a 2 sx halo, and gx.
__global__ __launch_bounds__(256) void wfr_curv_rows_kernel(const float* __restrict__ wx,
                                                            const float* __restrict__ wy,
                                                            const float* __restrict__ gx, int W,
                                                            int sx, float2* __restrict__ R) {
  extern __shared__ float lds[];
  const int span = 256 + 2 * sx;
  float* sa = lds;
  float* sb = lds + span;
  float* sg = lds + 2 * span;
  const int i = blockIdx.y, j0 = blockIdx.x * 256;
  const float* rwx = wx + (size_t)i * W;
  const float* rwy = wy + (size_t)i * W;
  for (int t = threadIdx.x; t < span; t += 256) {
    int jj = j0 - sx + t;
    jj = jj < 0 ? 0 : (jj >= W ? W - 1 : jj);
    sa[t] = rwx[jj];
    sb[t] = rwy[jj];
  }
  for (int t = threadIdx.x; t <= 2 * sx; t += 256) sg[t] = gx[t];
  __syncthreads();
  const int j = j0 + threadIdx.x;
  if (j >= W) return;
  float ax = 0.f, ay = 0.f;
  // output j reads wx(j - k) = sa[threadIdx.x + sx - k], k = -sx..sx
  for (int k = -sx; k <= sx; ++k) {
    const float g = sg[k + sx];
    const int o = threadIdx.x + sx - k;
    ax = fmaf((float)k * g, sa[o], ax);
    ay = fmaf(g, sb[o], ay);
  }
  R[(size_t)i * W + j] = make_float2(ax, ay);
}

// Column pass + compensation: cxx = -(sum_k gy(k) Rx(i-k, j)) / Sx,
// cyy = -(sum_k k gy(k) Ry(i-k, j)) / Sy (row index clamped), then
// phase_comp = wrap(phase - atan(sigx^2 cxx)/2 - atan(sigy^2 cyy)/2) and
// b = r ((1 + sigx^4 cxx^2)/(4 pi sigx^2))^(1/4) ((1 + sigy^4 cyy^2)/(4 pi sigy^2))^(1/4)
// (WFT/WFT2_CPU.cpp:674-691).  Tiles of 32 columns x 32 rows; 256 threads
// = 32 columns x 8 groups of 4 rows.  LDS: (32 + 2 sy) rows x 32 columns of R, and gy.
constexpr int kCurvTile = 32;
__global__ __launch_bounds__(256) void wfr_curv_cols_kernel(const float2* __restrict__ R,
                                                            const float* __restrict__ gy, int W,
                                                            int H, int sy, WfrConst c, WfrOut z) {
  extern __shared__ float2 lds2[];
  const int rows = kCurvTile + 2 * sy;
  float2* sR = lds2;                                            // [rows][32]
  float* sg = reinterpret_cast<float*>(lds2 + rows * kCurvTile);  // [2 sy + 1]
  const int j0 = blockIdx.x * kCurvTile, i0 = blockIdx.y * kCurvTile;
  for (int t = threadIdx.x; t < rows * kCurvTile; t += 256) {
    const int rr = t / kCurvTile, cc = t % kCurvTile;
    int ii = i0 - sy + rr;
    ii = ii < 0 ? 0 : (ii >= H ? H - 1 : ii);
    const int jj = j0 + cc < W ? j0 + cc : W - 1;
    sR[t] = R[(size_t)ii * W + jj];
  }
  for (int t = threadIdx.x; t <= 2 * sy; t += 256) sg[t] = gy[t];
  __syncthreads();
  const int cc = threadIdx.x % kCurvTile, grp = threadIdx.x / kCurvTile;
  const int j = j0 + cc;
  if (j >= W) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int li = grp * 4 + u, i = i0 + li;
    if (i >= H) break;
    float ax = 0.f, ay = 0.f;
    // output i reads R(i - k) = sR[(li + sy - k) * 32 + cc], k = -sy..sy
    for (int k = -sy; k <= sy; ++k) {
      const float g = sg[k + sy];
      const float2 v = sR[(li + sy - k) * kCurvTile + cc];
      ax = fmaf(g, v.x, ax);
      ay = fmaf((float)k * g, v.y, ay);
    }
    const float cxx = -ax * c.inv_sx, cyy = -ay * c.inv_sy;
    const size_t id = (size_t)i * W + j;
    z.cxx[id] = cxx;
    z.cyy[id] = cyy;
    float pc = z.phase[id] - 0.5f * atanf(c.sig2x * cxx) - 0.5f * atanf(c.sig2y * cyy);
    pc = atan2f(sinf(pc), cosf(pc));
    z.phase_comp[id] = pc;
    z.b[id] = z.r[id] *
              powf((1.f + c.sig2x * c.sig2x * cxx * cxx) * c.q4x, 0.25f) *
              powf((1.f + c.sig2y * c.sig2y * cyy * cyy) * c.q4y, 0.25f);
  }
}

hipError_t launch_wfr_cols(const WffPlan& p, const WfrBufs& w, hipStream_t s) {
  // columns: ridge over wy per (wx, column)
  const int waves = p.cols_waves;
  const int smem = plan_smem_bytes(p.plany, waves, true);
  switch (p.plany) {
#define BODY                                                                                 \
  {                                                                                          \
    const int per = items_per_block<P>(waves);                                               \
    const int grid = p.nwx * ((p.W + per - 1) / per);                                        \
    switch (waves) {                                                                         \
      case 1:                                                                                \
        hipLaunchKernelGGL((wfr_cols_kernel<P, 1>), dim3(grid), dim3(64), smem, s, p.d_A,   \
                           p.d_D, w.d_idx, p.d_Hy, p.d_twy, p.W, p.H, p.nwx, p.nwy,          \
                           p.plane_elems);                                                   \
        break;                                                                               \
      case 2:                                                                                \
        hipLaunchKernelGGL((wfr_cols_kernel<P, 2>), dim3(grid), dim3(128), smem, s, p.d_A,  \
                           p.d_D, w.d_idx, p.d_Hy, p.d_twy, p.W, p.H, p.nwx, p.nwy,          \
                           p.plane_elems);                                                   \
        break;                                                                               \
      default:                                                                               \
        hipLaunchKernelGGL((wfr_cols_kernel<P, 4>), dim3(grid), dim3(256), smem, s, p.d_A,  \
                           p.d_D, w.d_idx, p.d_Hy, p.d_twy, p.W, p.H, p.nwx, p.nwy,          \
                           p.plane_elems);                                                   \
        break;                                                                               \
    }                                                                                        \
  }
    GLSU_PLANS(GLSU_DISPATCH_CASE)
#undef BODY
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_wfr_post(const WffPlan& p, const WfrBufs& w, const WfrOut& z, hipStream_t s) {
  const int gr = (int)((p.plane_elems + 255) / 256);
  hipLaunchKernelGGL(wfr_ridge_kernel, dim3(gr), dim3(256), 0, s, p.d_D, w.d_idx, p.W, p.H, p.nwx,
                     p.plane_elems, p.wxl, p.wxi, p.wyl, p.wyi, z);
  // curvature: the A planes are free now and hold the row pass
  float2* R = p.d_A;
  hipLaunchKernelGGL(wfr_curv_rows_kernel, dim3((p.W + 255) / 256, p.H), dim3(256),
                     sizeof(float) * (2 * (256 + 2 * p.sx) + 2 * p.sx + 1), s, z.wx, z.wy, w.d_gx,
                     p.W, p.sx, R);
  const size_t sm = sizeof(float2) * (kCurvTile + 2 * p.sy) * kCurvTile +
                    sizeof(float) * (2 * p.sy + 1);
  hipLaunchKernelGGL(wfr_curv_cols_kernel,
                     dim3((p.W + kCurvTile - 1) / kCurvTile, (p.H + kCurvTile - 1) / kCurvTile),
                     dim3(256), sm, s, R, w.d_gy, p.W, p.H, p.sy, w.c, z);
  return hipGetLastError();
}

hipError_t launch_wfr(const WffPlan& p, const WfrBufs& w, const WfrOut& z, hipStream_t s) {
  hipError_t e = launch_wfr_cols(p, w, s);
  return e != hipSuccess ? e : launch_wfr_post(p, w, z, s);
}

size_t wfr_curv_lds_bytes(int sx, int sy) {
  const size_t rows = sizeof(float) * (2 * (256 + 2 * (size_t)sx) + 2 * sx + 1);
  const size_t cols = sizeof(float2) * (kCurvTile + 2 * (size_t)sy) * kCurvTile +
                      sizeof(float) * (2 * sy + 1);
  return rows > cols ? rows : cols;
}

}  // namespace glsu

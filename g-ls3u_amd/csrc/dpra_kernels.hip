// dpra_kernels.hip -- DPRA 3x3 least squares on gfx950 (the per-pixel
// solution is in dpra_device.h), the learned path's update and spot-check
// residual, and the synthetic fringe generator.
#include "dpra_device.h"

#include <algorithm>

namespace glsu {

// Pixel tiles of TX x TY (TX floats = one row segment), one pixel per thread
// (several pixels per thread measured slower: the per-pixel body branches, so
// their long dependency chains serialise), walked row-major and
// XCD-contiguous so that halo rows shared by neighbouring tiles are fetched
// from HBM once per L2.  The shape is a tuning choice (kLsTiles); the
// per-pixel arithmetic does not depend on it.  64 x 8 is the default and the
// batched kernel's shape.
constexpr int DTX = 64, DTY = 8;
constexpr int DTHREADS = DTX * DTY;

// One TX x TY tile (bx, by) of one image; all threads of the block take part.
template <typename InT, int TX, int TY>
__device__ __forceinline__ void dpra_ls_body(const InT* __restrict__ I,
                                             const float* __restrict__ phr,
                                             float2* __restrict__ z, int W, int H, int pitch,
                                             int bx, int by) {
  constexpr int NT = TX * TY;
  __shared__ float sI[TY + 2][TX + 2];
  __shared__ float sS[TY + 2][TX + 2];  // sin(phi/2)
  __shared__ float sC[TY + 2][TX + 2];  // cos(phi/2)
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x0 = bx * TX - 1, y0 = by * TY - 1;
  // Staging of the (TX+2) x (TY+2) tile with its halo: every thread stages
  // its own pixel (LDS (ty+1, tx+1)) and the first NH threads one halo pixel
  // each (top row, bottom row, left column, right column), so the slot
  // coordinates need no division.  Every load is issued before any is used,
  // from clamped (always in-bounds) coordinates, and the out-of-image values
  // are replaced afterwards: loads under per-slot guards would each wait for
  // the previous slot's round trip.
  constexpr int NH = 2 * (TX + 2) + 2 * TY;
  static_assert(NH <= NT, "at most one halo pixel per thread");
  const int h = threadIdx.x;
  const bool halo = h < NH;
  int hy, hx;  // halo slot of thread h (h < NH)
  if (h < TX + 2) {
    hy = 0; hx = h;
  } else if (h < 2 * (TX + 2)) {
    hy = TY + 1; hx = h - (TX + 2);
  } else if (h < 2 * (TX + 2) + TY) {
    hy = h - 2 * (TX + 2) + 1; hx = 0;
  } else {
    hy = min(h - 2 * (TX + 2) - TY + 1, TY); hx = TX + 1;
  }
  const int ly[2] = {ty + 1, hy}, lx[2] = {tx + 1, hx};
  float ld_i[2], ld_p[2];
  bool in[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int gx = x0 + lx[k], gy = y0 + ly[k];
    in[k] = gx >= 0 && gx < W && gy >= 0 && gy < H;
    const int cx = min(max(gx, 0), W - 1), cy = min(max(gy, 0), H - 1);
    if (k == 0 || halo) {
      ld_i[k] = load_in(I, (size_t)cy * pitch + cx);
      ld_p[k] = phr[(size_t)cy * W + cx];
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k == 1 && !halo) break;
    float hs = 0.f, hc = 1.f;
    sincos_half(0.5f * ld_p[k], &hs, &hc);
    sI[ly[k]][lx[k]] = in[k] ? ld_i[k] : 0.f;
    sS[ly[k]][lx[k]] = in[k] ? hs : 0.f;
    sC[ly[k]][lx[k]] = in[k] ? hc : 1.f;
  }
  __syncthreads();
  const int x = bx * TX + tx, y = by * TY + ty;
  if (x >= W || y >= H) return;
  z[(size_t)y * W + x] =
      dpra_pixel(&sI[0][0], &sS[0][0], &sC[0][0], TX + 2, ty + 1, tx + 1, x, y, W, H);
}

template <typename InT, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void dpra_ls_kernel(const InT* __restrict__ I,
                                                          const float* __restrict__ phr,
                                                          float2* __restrict__ z, int W, int H,
                                                          int pitch) {
  const int nbx = (W + TX - 1) / TX;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  dpra_ls_body<InT, TX, TY>(I, phr, z, W, H, pitch, tile % nbx, tile / nbx);
}

// Batched form: persistent walk over the 64 x 8 tiles of every image in the
// table; frame t of image s is frames_s + t*W*H, its reference phase the
// buffer that the swap parity of t names.
template <typename InT>
__global__ __launch_bounds__(DTHREADS) void dpra_ls_batch_kernel(const BatchSeg* __restrict__ tab,
                                                                 int nseg, int ntiles,
                                                                 const int* __restrict__ frame) {
  const int t = *frame;
  for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
    const int tile = xcd_remap(v, ntiles);
    int lo = 0, hi = nseg - 1;  // last image whose first tile is <= tile
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[mid].tile0 <= tile) lo = mid;
      else hi = mid - 1;
    }
    const BatchSeg& s = tab[lo];
    const int nbx = (s.W + DTX - 1) / DTX;
    const int lt = tile - s.tile0;
    const int c = (t == 0 ? 0 : (t - 1) / s.rr + 1) & 1;
    __syncthreads();  // LDS of the previous tile is free
    dpra_ls_body<InT, DTX, DTY>((const InT*)s.frames + (size_t)t * s.W * s.H, s.phi[c], s.z, s.W,
                                s.H, s.W, lt % nbx, lt / nbx);
  }
}

hipError_t launch_dpra_ls_batch(const BatchSeg* tab, int nseg, int ntiles, int grid, int dtype,
                                const int* frame, hipStream_t s) {
  switch (dtype) {
    case IN_U8:
      hipLaunchKernelGGL(dpra_ls_batch_kernel<uint8_t>, dim3(grid), dim3(DTHREADS), 0, s, tab,
                         nseg, ntiles, frame);
      break;
    case IN_F32:
      hipLaunchKernelGGL(dpra_ls_batch_kernel<float>, dim3(grid), dim3(DTHREADS), 0, s, tab,
                         nseg, ntiles, frame);
      break;
    case IN_F64:
      hipLaunchKernelGGL(dpra_ls_batch_kernel<double>, dim3(grid), dim3(DTHREADS), 0, s, tab,
                         nseg, ntiles, frame);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Learned path: phase update from the network output, and the tile residual.
__global__ __launch_bounds__(256) void net_update_kernel(const float2* __restrict__ out,
                                                         UpdateArgs u, size_t n,
                                                         const int* __restrict__ skip) {
  if (skip && *skip) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 o = out[i];
  phase_update(u, i, atan2f(o.y, o.x));
}

hipError_t launch_net_update(const float2* out, const UpdateArgs& u, size_t n, hipStream_t s,
                             const int* skip) {
  hipLaunchKernelGGL(net_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out,
                     u, n, skip);
  return hipGetLastError();
}

// RMS over a 64 x 64 tile of wrap(angle(net) - angle(fwin)); sets the
// fallback flag (and its complement) and accumulates the statistics
__global__ __launch_bounds__(1024) void tile_residual_kernel(
    const float2* __restrict__ net, int W, int tx, int ty, const float2* __restrict__ fwin, int Ww,
    int ox, int oy, float tol, int* __restrict__ flag, int* __restrict__ not_flag,
    float* __restrict__ stats) {
  // 4096 pixels, four per thread: all eight loads are issued before the first
  // atan2, then the squares are summed per wave (lane exchange) and across
  // the 16 waves in LDS
  __shared__ float red[16];
  float2 a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = threadIdx.x + 1024 * k, r = i / 64, c = i % 64;
    a[k] = net[(size_t)(ty + r) * W + tx + c];
    b[k] = fwin[(size_t)(oy + r) * Ww + ox + c];
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // angle(a conj(b)) = wrapped phase difference
    const float d = atan2f(a[k].y * b[k].x - a[k].x * b[k].y, a[k].x * b[k].x + a[k].y * b[k].y);
    s += d * d;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += red[w];
    const float rms = sqrtf(t / 4096.f);
    const int fb = rms > tol ? 1 : 0;
    *flag = fb;
    *not_flag = 1 - fb;
    stats[0] += 1.f;
    stats[1] += (float)fb;
    stats[2] += rms;
    stats[3] = fmaxf(stats[3], rms);
    stats[4] = rms;  // last frame's value
  }
}

hipError_t launch_tile_residual(const float2* net, int W, int tx, int ty, const float2* fwin,
                                int Ww, int ox, int oy, float tol, int* flag, int* not_flag,
                                float* stats, hipStream_t s) {
  hipLaunchKernelGGL(tile_residual_kernel, dim3(1), dim3(1024), 0, s, net, W, tx, ty, fwin, Ww, ox,
                     oy, tol, flag, not_flag, stats);
  return hipGetLastError();
}

int batch_grid_ls(int ntiles) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, dpra_ls_batch_kernel<float>, DTHREADS,
                                                    0) != hipSuccess)
    return -1;
  const int cap = std::max(8, (cus * std::max(per, 1)) / 8 * 8);
  return ntiles <= cap ? ntiles : cap;
}

int dpra_ls_tiles(int W, int H) { return ((W + DTX - 1) / DTX) * ((H + DTY - 1) / DTY); }

// Least-squares tile shapes (tuning ids 1..kLsTileCount; 1 = 64 x 8).
#define GLSU_LS_TILES(X) X(1, 64, 8) X(2, 64, 4) X(3, 32, 8) X(4, 128, 4) X(5, 64, 16) X(6, 32, 16)

bool ls_tile_valid(int id) {
  switch (id) {
#define GLSU_LS_VALID(i, tx, ty) case i:
    GLSU_LS_TILES(GLSU_LS_VALID)
#undef GLSU_LS_VALID
    return true;
    default:
      return false;
  }
}

template <int TX, int TY>
static hipError_t launch_ls_shape(const void* frame, int dtype, const float* phiref, float2* z,
                                  int W, int H, int pitch, hipStream_t s) {
  dim3 grid(((W + TX - 1) / TX) * ((H + TY - 1) / TY)), block(TX * TY);
  switch (dtype) {
    case IN_U8:
      hipLaunchKernelGGL((dpra_ls_kernel<uint8_t, TX, TY>), grid, block, 0, s,
                         (const uint8_t*)frame, phiref, z, W, H, pitch);
      break;
    case IN_F32:
      hipLaunchKernelGGL((dpra_ls_kernel<float, TX, TY>), grid, block, 0, s, (const float*)frame,
                         phiref, z, W, H, pitch);
      break;
    case IN_F64:
      hipLaunchKernelGGL((dpra_ls_kernel<double, TX, TY>), grid, block, 0, s,
                         (const double*)frame, phiref, z, W, H, pitch);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dpra_ls(const void* frame, int dtype, const float* phiref, float2* z, int W,
                          int H, int pitch, hipStream_t s, int tile) {
  switch (tile) {
#define GLSU_LS_CASE(i, tx, ty) \
  case i:                       \
    return launch_ls_shape<tx, ty>(frame, dtype, phiref, z, W, H, pitch, s);
    GLSU_LS_TILES(GLSU_LS_CASE)
#undef GLSU_LS_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Synthetic fringes: I = 128 + 100 cos(phi) + noise * N(0,1),
// phi = (amp0 + amp1 t) peaks(X, Y) + 2 pi carrier x, X/Y = linspace(-3, 3).
// N(0,1): SplitMix64 chained over (seed, t, pixel) + FP64 Box-Muller (cosine).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t t, uint64_t p) {
  return mix64(mix64(mix64(seed) ^ t) ^ p);
}
__device__ __forceinline__ double peaks(double X, double Y) {
  return 3.0 * (1.0 - X) * (1.0 - X) * exp(-X * X - (Y + 1.0) * (Y + 1.0)) -
         10.0 * (X / 5.0 - X * X * X - Y * Y * Y * Y * Y) * exp(-X * X - Y * Y) -
         exp(-(X + 1.0) * (X + 1.0) - Y * Y) / 3.0;
}

__global__ __launch_bounds__(256) void gen_frames_kernel(float* __restrict__ out, int W, int H,
                                                         int t0, int T, uint64_t seed,
                                                         double amp0, double amp1,
                                                         double carrier, double noise) {
  const size_t n = (size_t)W * H;
  const size_t total = n * T;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (size_t)gridDim.x * blockDim.x) {
    const int tt = (int)(g / n);
    const size_t p = g % n;
    const int i = (int)(p / W), j = (int)(p % W);
    const double X = W > 1 ? -3.0 + 6.0 * (double)j / (double)(W - 1) : 3.0;
    const double Y = H > 1 ? -3.0 + 6.0 * (double)i / (double)(H - 1) : 3.0;
    const int t = t0 + tt;
    const double amp = amp0 + amp1 * (double)t;
    const double ph = amp * peaks(X, Y) + 2.0 * M_PI * carrier * (double)j;
    const uint64_t h1 = hash3(seed, (uint64_t)t, 2 * p), h2 = hash3(seed, (uint64_t)t, 2 * p + 1);
    const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    const double nz = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    out[g] = (float)(128.0 + 100.0 * cos(ph) + noise * nz);
  }
}

hipError_t launch_gen_frames(float* out, int W, int H, int t0, int T, uint64_t seed, double amp0,
                             double amp1, double carrier, double noise, hipStream_t s) {
  hipLaunchKernelGGL(gen_frames_kernel, dim3(4096), dim3(256), 0, s, out, W, H, t0, T, seed, amp0,
                     amp1, carrier, noise);
  return hipGetLastError();
}

}  // namespace glsu

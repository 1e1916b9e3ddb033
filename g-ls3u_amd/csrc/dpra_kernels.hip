// dpra_kernels.hip -- DPRA 3x3 least squares on gfx950, and the synthetic
// fringe generator.
//
// Reference (DPRA/dpra_cpu.cpp:153-232): for every pixel p, over the clipped
// 3x3 neighbourhood N(p) with c = cos(phiref), s = sin(phiref),
//   A = [[9, Sc, Ss], [Sc, Scc, Scs], [Ss, Scs, Sss]]   (A00 = 9 even at borders)
//   b = [SI, SIc, SIs],  x = A^-1 b,  Delta = atan2(-x2, x1),  z = e^{j Delta}.
//
// Same solution, better conditioned arithmetic.  Rotate the (c, s) basis by
// the centre phase phi_p and shift the cosine column by one:
//   w_q = [1, e_q, s'_q],  e_q = cos(d_q) - 1 = -2 sin^2(d_q/2),  s'_q = sin(d_q),
//   d_q = phi_q - phi_p.
// A = R^T M A'' M^T R with A'' = sum_q w_q w_q^T + (9 - n) w0 w0^T, w0 = [1,-1,0],
// and b'' = sum_q I_q w_q, so A'' y = b'' gives x'_1 = y1, x'_2 = y2 and
//   Delta = atan2(-y2, y1) - phi_p,  i.e.  z = (y1 - j y2)/|y| * e^{-j phi_p}.
// The half-angle sines come from one sincos(phi/2) per pixel and the
// difference formula sin(d/2) = sin(a/2)cos(b/2) - cos(a/2)sin(b/2), which has
// no cancellation for small d, so e and s' keep full relative precision.  The
// three sums whose terms change sign (s', e s', I s') are accumulated with a
// compensated two-sum; the other five have terms of one sign and are plain
// FP32 sums.  The DC unknown is eliminated (2x2 Schur complement) and the 2x2
// system is solved by Cramer's rule; det(A) = 9 det2.  Where
// |det A| < 1e-6 * A00*A11*A22 the 2x2 solve is redone in FP64 from the
// compensated sums; where det A <= 1e-10 * A00*A11*A22 the pixel is singular
// and z = 1 (the oracle's defined behaviour for that case).
#include "glsu_internal.h"

#include <algorithm>

#include <cstdint>

namespace glsu {

struct Comp {  // compensated FP32 accumulator (Knuth TwoSum)
  float s, c;
  __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
  __device__ __forceinline__ void add(float x) {
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ __forceinline__ float f() const { return s + c; }
  __device__ __forceinline__ double d() const { return (double)s + (double)c; }
};

template <typename InT>
__device__ __forceinline__ float load_in(const InT* p, size_t i) {
  return (float)p[i];
}

// sin and cos of x for |x| <= pi/2 (half of a wrapped phase): Taylor series
// to x^13 / x^14 in Horner form; truncation error < 7e-10, below FP32 rounding.
__device__ __forceinline__ void sincos_half(float x, float* s, float* c) {
  const float x2 = x * x;
  float ps = -1.f / 6227020800.f;              // -1/13!
  ps = fmaf(ps, x2, 1.f / 39916800.f);
  ps = fmaf(ps, x2, -1.f / 362880.f);
  ps = fmaf(ps, x2, 1.f / 5040.f);
  ps = fmaf(ps, x2, -1.f / 120.f);
  ps = fmaf(ps, x2, 1.f / 6.f);
  *s = fmaf(-x * x2, ps, x);                   // x - x^3 (1/3! - x^2/5! + ...)
  float pc = 1.f / 87178291200.f;              // 1/14!
  pc = fmaf(pc, x2, -1.f / 479001600.f);
  pc = fmaf(pc, x2, 1.f / 3628800.f);
  pc = fmaf(pc, x2, -1.f / 40320.f);
  pc = fmaf(pc, x2, 1.f / 720.f);
  pc = fmaf(pc, x2, -1.f / 24.f);
  pc = fmaf(pc, x2, 0.5f);
  *c = fmaf(-x2, pc, 1.f);                     // 1 - x^2 (1/2! - x^2/4! + ...)
}

// 64 x 8 pixel tiles (256-byte rows), one pixel per thread (DRPT rows per
// thread is kept as a parameter: the per-pixel body branches, so several
// pixels per thread serialise their long dependency chains and measured
// slower), walked row-major and XCD-contiguous so that halo rows shared by
// neighbouring tiles are fetched from HBM once per L2.
constexpr int DTX = 64, DTY = 8, DRPT = 1;
constexpr int DTHREADS = DTX * DTY / DRPT;

// One 64 x 16 tile (bx, by) of one image; all threads of the block take part.
template <typename InT>
__device__ __forceinline__ void dpra_ls_body(const InT* __restrict__ I,
                                             const float* __restrict__ phr,
                                             float2* __restrict__ z, int W, int H, int pitch,
                                             int bx, int by) {
  __shared__ float sI[DTY + 2][DTX + 2];
  __shared__ float sS[DTY + 2][DTX + 2];  // sin(phi/2)
  __shared__ float sC[DTY + 2][DTX + 2];  // cos(phi/2)
  const int tx = threadIdx.x % DTX, ty0 = (threadIdx.x / DTX) * DRPT;
  const int x0 = bx * DTX - 1, y0 = by * DTY - 1;
  for (int i = threadIdx.x; i < (DTX + 2) * (DTY + 2); i += DTHREADS) {
    const int ly = i / (DTX + 2), lx = i % (DTX + 2);
    const int gx = x0 + lx, gy = y0 + ly;
    float vi = 0.f, hs = 0.f, hc = 1.f;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
      vi = load_in(I, (size_t)gy * pitch + gx);
      sincos_half(0.5f * phr[(size_t)gy * W + gx], &hs, &hc);
    }
    sI[ly][lx] = vi;
    sS[ly][lx] = hs;
    sC[ly][lx] = hc;
  }
  __syncthreads();
  const int x = bx * DTX + tx;
  if (x >= W) return;
#pragma unroll
  for (int r = 0; r < DRPT; ++r) {
  const int ty = ty0 + r, y = by * DTY + ty;
  if (y >= H) continue;

  const float hsp = sS[ty + 1][tx + 1], hcp = sC[ty + 1][tx + 1];
  float Se = 0.f, See = 0.f, Sss = 0.f, SI = 0.f, SIe = 0.f;
  Comp Ss, Ses, SIs;
  Ss.init(); Ses.init(); SIs.init();
  int n = 0;
  auto neighbour = [&](int dy, int dx) {
    const float hsq = sS[ty + 1 + dy][tx + 1 + dx], hcq = sC[ty + 1 + dy][tx + 1 + dx];
    // sin/cos of (phi_q - phi_p)/2 by the difference formulas
    const float sh = fmaf(hsq, hcp, -hcq * hsp);
    const float ch = fmaf(hcq, hcp, hsq * hsp);
    const float sq = 2.f * sh * ch;
    const float e = -2.f * sh * sh;
    const float Iq = sI[ty + 1 + dy][tx + 1 + dx];
    Se += e; See += e * e; Sss += sq * sq; SI += Iq; SIe += Iq * e;
    Ss.add(sq); Ses.add(e * sq); SIs.add(Iq * sq);
  };
  if (x > 0 && x < W - 1 && y > 0 && y < H - 1) {  // interior: full neighbourhood
    n = 9;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) neighbour(dy, dx);
  } else {  // image border: clipped neighbourhood (4 or 6 pixels)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int gx = x + dx, gy = y + dy;
        if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
        ++n;
        neighbour(dy, dx);
      }
  }
  // cos/sin of phi_p by the double-angle formulas
  const float cp = hcp * hcp - hsp * hsp, sp = 2.f * hsp * hcp;
  const float q = (float)(9 - n);
  // diagonal of the unrotated A: A11 = sum c^2, A22 = sum s^2
  const float s1e2 = (float)n + 2.f * Se + See;   // sum (1+e)^2
  const float s1es = Ss.f() + Ses.f();             // sum (1+e) s'
  const float A11 = cp * cp * s1e2 - 2.f * cp * sp * s1es + sp * sp * Sss;
  const float A22 = sp * sp * s1e2 + 2.f * cp * sp * s1es + cp * cp * Sss;
  const float dg = 9.f * A11 * A22;

  // FP32 Schur complement of A'' on the DC unknown
  const float a01 = Se - q, a02 = Ss.f(), a11 = See + q, a12 = Ses.f(), a22 = Sss;
  const float b0 = SI, b1 = SIe, b2 = SIs.f();
  const float s11 = a11 - a01 * a01 * (1.f / 9.f);
  const float s12 = a12 - a01 * a02 * (1.f / 9.f);
  const float s22 = a22 - a02 * a02 * (1.f / 9.f);
  const float r1 = b1 - a01 * b0 * (1.f / 9.f);
  const float r2 = b2 - a02 * b0 * (1.f / 9.f);
  float det2 = s11 * s22 - s12 * s12;
  float y1 = r1 * s22 - s12 * r2;
  float y2 = s11 * r2 - s12 * r1;
  bool singular = false;
  if (!(fabsf(9.f * det2) >= 1e-6f * fabsf(dg))) {
    // near-singular: FP64 solve from the compensated sums
    const double qd = (double)(9 - n);
    const double A01 = (double)Se - qd, A02 = Ss.d(), A11d = (double)See + qd, A12 = Ses.d(),
                 A22d = (double)Sss;
    const double B0 = SI, B1 = SIe, B2 = SIs.d();
    const double S11 = A11d - A01 * A01 / 9.0, S12 = A12 - A01 * A02 / 9.0,
                 S22 = A22d - A02 * A02 / 9.0;
    const double R1 = B1 - A01 * B0 / 9.0, R2 = B2 - A02 * B0 / 9.0;
    const double D2 = S11 * S22 - S12 * S12;
    const double Y1 = R1 * S22 - S12 * R2, Y2 = S11 * R2 - S12 * R1;
    if (!(fabs(9.0 * D2) > 1e-10 * fabs((double)dg))) {
      singular = true;
    } else {
      const double sg = D2 < 0 ? -1.0 : 1.0;
      const double nr = sqrt(Y1 * Y1 + Y2 * Y2);
      y1 = (float)(sg * Y1 / nr);
      y2 = (float)(sg * Y2 / nr);
      det2 = 1.f;
    }
  }
  float2 out;
  if (singular || (y1 == 0.f && y2 == 0.f)) {
    out = make_float2(1.f, 0.f);
  } else {
    if (det2 < 0.f) { y1 = -y1; y2 = -y2; }
    const float inv = rsqrtf(y1 * y1 + y2 * y2);
    const float ur = y1 * inv, ui = -y2 * inv;      // e^{j theta'}
    out = make_float2(ur * cp + ui * sp, ui * cp - ur * sp);  // * e^{-j phi_p}
  }
  z[(size_t)y * W + x] = out;
  }  // rows of this thread
}

template <typename InT>
__global__ __launch_bounds__(DTHREADS) void dpra_ls_kernel(const InT* __restrict__ I,
                                                           const float* __restrict__ phr,
                                                           float2* __restrict__ z, int W, int H,
                                                           int pitch) {
  const int nbx = (W + DTX - 1) / DTX;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  dpra_ls_body<InT>(I, phr, z, W, H, pitch, tile % nbx, tile / nbx);
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
    dpra_ls_body<InT>((const InT*)s.frames + (size_t)t * s.W * s.H, s.phi[c], s.z, s.W, s.H, s.W,
                      lt % nbx, lt / nbx);
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
__global__ __launch_bounds__(256) void tile_residual_kernel(
    const float2* __restrict__ net, int W, int tx, int ty, const float2* __restrict__ fwin, int Ww,
    int ox, int oy, float tol, int* __restrict__ flag, int* __restrict__ not_flag,
    float* __restrict__ stats) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int r = i / 64, c = i % 64;
    const float2 a = net[(size_t)(ty + r) * W + tx + c];
    const float2 b = fwin[(size_t)(oy + r) * Ww + ox + c];
    // angle(a conj(b)) = wrapped phase difference
    const float d = atan2f(a.y * b.x - a.x * b.y, a.x * b.x + a.y * b.y);
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float rms = sqrtf(red[0] / 4096.f);
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
  hipLaunchKernelGGL(tile_residual_kernel, dim3(1), dim3(256), 0, s, net, W, tx, ty, fwin, Ww, ox,
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

hipError_t launch_dpra_ls(const void* frame, int dtype, const float* phiref, float2* z, int W,
                          int H, int pitch, hipStream_t s) {
  dim3 grid(((W + DTX - 1) / DTX) * ((H + DTY - 1) / DTY)), block(DTHREADS);
  switch (dtype) {
    case IN_U8:
      hipLaunchKernelGGL(dpra_ls_kernel<uint8_t>, grid, block, 0, s, (const uint8_t*)frame,
                         phiref, z, W, H, pitch);
      break;
    case IN_F32:
      hipLaunchKernelGGL(dpra_ls_kernel<float>, grid, block, 0, s, (const float*)frame, phiref,
                         z, W, H, pitch);
      break;
    case IN_F64:
      hipLaunchKernelGGL(dpra_ls_kernel<double>, grid, block, 0, s, (const double*)frame,
                         phiref, z, W, H, pitch);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
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

// dpra_device.h -- per-pixel DPRA 3x3 least squares (device code shared by the
// single-image and batched least-squares kernels).
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
// three sums whose terms change sign (s', e s', I s') are accumulated in FP64
// (exact for nine FP32 terms; FP64 adds and conversions issue at the FP32
// rate on this target, cheaper than a compensated FP32 two-sum); the other
// five have terms of one sign and are plain FP32 sums.  The DC unknown is eliminated (2x2 Schur complement) and the 2x2
// system is solved by Cramer's rule; det(A) = 9 det2.  Where
// |det A| < 1e-6 * A00*A11*A22 the 2x2 solve is redone in FP64 from the
// compensated sums; where det A <= 1e-10 * A00*A11*A22 the pixel is singular
// and z = 1 (the oracle's defined behaviour for that case).
#pragma once
#include "glsu_internal.h"

#include <cstdint>

namespace glsu {

struct Acc64 {  // FP64 accumulator of FP32 terms (exact for these 9-term sums)
  double s;
  __device__ __forceinline__ void init() { s = 0.0; }
  __device__ __forceinline__ void add(float x) { s += (double)x; }
  __device__ __forceinline__ float f() const { return (float)s; }
  __device__ __forceinline__ double d() const { return s; }
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

// The least-squares solution at one pixel from LDS stagings of I, sin(phi/2)
// and cos(phi/2) with row pitch `pitch`: (ly, lx) is the pixel's position in
// the staging (its neighbours at ly +- 1, lx +- 1), (x, y) its image position
// (the border test and the clipped neighbourhood use the image size W x H).
__device__ __forceinline__ float2 dpra_pixel(const float* sI, const float* sS, const float* sC,
                                             int pitch, int ly, int lx, int x, int y, int W,
                                             int H) {
  const float hsp = sS[ly * pitch + lx], hcp = sC[ly * pitch + lx];
  float Se = 0.f, See = 0.f, Sss = 0.f, SI = 0.f, SIe = 0.f;
  Acc64 Ss, Ses, SIs;
  Ss.init(); Ses.init(); SIs.init();
  int n = 0;
  auto neighbour = [&](int dy, int dx) {
    const float hsq = sS[(ly + dy) * pitch + lx + dx], hcq = sC[(ly + dy) * pitch + lx + dx];
    // sin/cos of (phi_q - phi_p)/2 by the difference formulas
    const float sh = fmaf(hsq, hcp, -hcq * hsp);
    const float ch = fmaf(hcq, hcp, hsq * hsp);
    const float sq = 2.f * sh * ch;
    const float e = -2.f * sh * sh;
    const float Iq = sI[(ly + dy) * pitch + lx + dx];
    Se += e; See += e * e; Sss += sq * sq; SI += Iq; SIe += Iq * e;
    Ss.add(sq); Ses.add(e * sq); SIs.add(Iq * sq);
  };
  // the centre's own term has d = 0 exactly (e = s' = 0): it adds I_p to SI and nothing
  // else (through the difference formula it would add the rounding residue of sh)
  auto centre = [&]() { SI += sI[ly * pitch + lx]; };
  if (x > 0 && x < W - 1 && y > 0 && y < H - 1) {  // interior: full neighbourhood
    n = 9;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        if (dy == 0 && dx == 0) centre();
        else neighbour(dy, dx);
      }
  } else {  // image border: clipped neighbourhood (4 or 6 pixels)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int gx = x + dx, gy = y + dy;
        if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
        ++n;
        if (dy == 0 && dx == 0) centre();
        else neighbour(dy, dx);
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
  return out;
}

}  // namespace glsu

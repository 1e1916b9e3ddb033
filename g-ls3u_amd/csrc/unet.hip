// unet.hip -- the learned WFF2 surrogate: a U-Net (levels x base channels,
// 3x3 convs, relu, max pooling, nearest x2 upsampling with skip
// concatenation, 1x1 output with a residual on the input) trained and run by
// the kernels of unet_kernels.hip.  BF16 weights and activations, FP32
// accumulation, FP32 master weights and Adam moments.
//
// The network has no counterpart in the reference codebase; its definition
// and gradients are those of the FP64 CPU reference used by the tests.
#include "../../include/glsu.h"
#include "glsu_internal.h"
#include "unet_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

namespace glsu {
void set_error(const char* msg);
}

using namespace glsu;

namespace {
struct Layer {
  int cin = 0, cout = 0, k = 3;       // k: 3 (3x3) or 1 (1x1)
  float *w = nullptr, *b = nullptr;   // FP32 master [cout][k][k][cin], [cout]
  float *gw = nullptr, *gb = nullptr;
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
  __bf16 *wf = nullptr, *wt = nullptr;  // BF16 forward / flipped-transposed copies
  long nw() const { return (long)cout * k * k * cin; }
};
struct Acts {  // activations and gradients for one (n, H, W)
  int n = 0, H = 0, W = 0;
  std::vector<void*> bufs;
  __bf16 *ea[8], *e[8], *p[8], *da[8], *d[8];
  __bf16 *ge[8], *gea[8], *gp[8], *gd[8], *gda[8], *gcat[8];
  float2* out = nullptr;
  float* loss = nullptr;
};
}  // namespace

struct glsu_unet {
  int levels = 4, base = 32;
  hipStream_t s = nullptr;
  std::vector<Layer> L;  // e0a e0b e1a e1b ... e{L-1}b d{L-2}a d{L-2}b ... d0a d0b out
  Acts a;
  int step = 0;
  int C(int l) const { return base << l; }
  Layer& enc(int l, int j) { return L[2 * l + j]; }
  Layer& dec(int l, int j) { return L[2 * levels + 2 * (levels - 2 - l) + j]; }
  Layer& out() { return L.back(); }
};

static int ufail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  set_error(buf);
  return code;
}
#define UTRY(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return ufail(GLSU_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static void acts_free(Acts& a) {
  for (void* p : a.bufs) hipFree(p);
  a.bufs.clear();
  a = Acts{};
}

static int acts_alloc(glsu_unet* h, int n, int H, int W) {
  Acts& a = h->a;
  if (a.n == n && a.H == H && a.W == W) return GLSU_OK;
  acts_free(a);
  a.n = n; a.H = H; a.W = W;
  auto get = [&](size_t elems, size_t sz) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, elems * sz) != hipSuccess) return nullptr;
    a.bufs.push_back(p);
    return p;
  };
  const int Lv = h->levels;
  for (int l = 0; l < Lv; ++l) {
    const size_t px = (size_t)n * (H >> l) * (W >> l), c = h->C(l);
    a.ea[l] = (__bf16*)get(px * c, 2);
    a.e[l] = (__bf16*)get(px * c, 2);
    a.ge[l] = (__bf16*)get(px * c, 2);
    a.gea[l] = (__bf16*)get(px * c, 2);
    if (l < Lv - 1) {
      const size_t pxl = px / 4;
      a.p[l] = (__bf16*)get(pxl * c, 2);
      a.gp[l] = (__bf16*)get(pxl * c, 2);
      a.da[l] = (__bf16*)get(px * c, 2);
      a.d[l] = (__bf16*)get(px * c, 2);
      a.gd[l] = (__bf16*)get(px * c, 2);
      a.gda[l] = (__bf16*)get(px * c, 2);
      a.gcat[l] = (__bf16*)get(px * (c + h->C(l + 1)), 2);
      if (!a.p[l] || !a.gp[l] || !a.da[l] || !a.d[l] || !a.gd[l] || !a.gda[l] || !a.gcat[l])
        return ufail(GLSU_ERR_DEVICE, "activation allocation failed");
    }
    if (!a.ea[l] || !a.e[l] || !a.ge[l] || !a.gea[l])
      return ufail(GLSU_ERR_DEVICE, "activation allocation failed");
  }
  a.out = (float2*)get((size_t)n * H * W, sizeof(float2));
  a.loss = (float*)get(1, sizeof(float));
  if (!a.out || !a.loss) return ufail(GLSU_ERR_DEVICE, "activation allocation failed");
  return GLSU_OK;
}

static int refresh_bf16(glsu_unet* h) {
  for (auto& l : h->L)
    if (l.wf) UTRY(unet_weights_bf16(l.w, l.wf, l.wt, l.cout, l.cin, h->s));
  return GLSU_OK;
}

static int check_shape(const glsu_unet* h, int n, int H, int W) {
  const int m = 1 << (h->levels - 1);
  if (n < 1 || H < 4 * m || W < 4 * m || H % (4 * m) || W % m)
    return ufail(GLSU_ERR_ARG,
                 "image %dx%d: height must be a multiple of %d and width of %d (levels %d)", W,
                 H, 4 * m, m, h->levels);
  return GLSU_OK;
}

// forward pass; x, out: n*H*W float2
static int forward(glsu_unet* h, const float2* x, float2* out, int n, int H, int W) {
  Acts& a = h->a;
  const int Lv = h->levels;
  hipStream_t s = h->s;
  Layer& e0a = h->enc(0, 0);
  UTRY(unet_conv3_in(x, e0a.w, e0a.b, a.ea[0], n, H, W, h->C(0), s));
  for (int l = 0; l < Lv; ++l) {
    const int Hl = H >> l, Wl = W >> l;
    if (l > 0) {  // a.p[l - 1]: pooled by the previous level's second conv
      Layer& la = h->enc(l, 0);
      UTRY(unet_conv3(a.p[l - 1], h->C(l - 1), nullptr, 0, la.wf, la.b, nullptr, a.ea[l], n, Hl,
                      Wl, h->C(l), 0, s));
    }
    Layer& lb = h->enc(l, 1);
    UTRY(unet_conv3(a.ea[l], h->C(l), nullptr, 0, lb.wf, lb.b, nullptr, a.e[l], n, Hl, Wl,
                    h->C(l), 0, s, l + 1 < Lv ? a.p[l] : nullptr));
  }
  const __bf16* low = a.e[Lv - 1];
  for (int l = Lv - 2; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l;
    Layer& la = h->dec(l, 0);
    Layer& lb = h->dec(l, 1);
    UTRY(unet_conv3(a.e[l], h->C(l), low, h->C(l + 1), la.wf, la.b, nullptr, a.da[l], n, Hl, Wl,
                    h->C(l), 0, s));
    UTRY(unet_conv3(a.da[l], h->C(l), nullptr, 0, lb.wf, lb.b, nullptr, a.d[l], n, Hl, Wl,
                    h->C(l), 0, s));
    low = a.d[l];
  }
  Layer& o = h->out();
  UTRY(unet_conv1_out(a.d[0], x, o.w, o.b, out, (long)n * H * W, h->C(0), s));
  return GLSU_OK;
}

// loss and gradients (accumulated into zeroed gw/gb) for targets u
static int backward(glsu_unet* h, const float2* x, const float2* u, int n, int H, int W) {
  Acts& a = h->a;
  const int Lv = h->levels;
  hipStream_t s = h->s;
  for (auto& l : h->L) {
    UTRY(unet_fill(l.gw, l.nw(), 0.f, s));
    UTRY(unet_fill(l.gb, l.cout, 0.f, s));
  }
  UTRY(unet_fill(a.loss, 1, 0.f, s));
  const long P = (long)n * H * W;
  Layer& o = h->out();
  UTRY(unet_conv1_out_bwd(a.d[0], a.out, u, o.w, a.gd[0], o.gw, o.gb, a.loss, P, h->C(0), s));
  // decoder, top level first; gd[l]: gradient at the pre-activation of d_l
  __bf16* ge_top = nullptr;
  for (int l = 0; l <= Lv - 2; ++l) {
    const int Hl = H >> l, Wl = W >> l;
    const long Pl = (long)n * Hl * Wl;
    const int Cl = h->C(l), Cn = h->C(l + 1);
    Layer& la = h->dec(l, 0);
    Layer& lb = h->dec(l, 1);
    const __bf16* low = (l + 1 <= Lv - 2) ? a.d[l + 1] : a.e[Lv - 1];
    UTRY(unet_conv3_dw(a.gd[l], Cl, a.da[l], Cl, nullptr, 0, lb.gw, n, Hl, Wl, s));
    UTRY(unet_bias_grad(a.gd[l], lb.gb, Pl, Cl, s));
    UTRY(unet_conv3(a.gd[l], Cl, nullptr, 0, lb.wt, nullptr, a.da[l], a.gda[l], n, Hl, Wl, Cl, 1,
                    s));
    UTRY(unet_conv3_dw(a.gda[l], Cl, a.e[l], Cl, low, Cn, la.gw, n, Hl, Wl, s));
    UTRY(unet_bias_grad(a.gda[l], la.gb, Pl, Cl, s));
    UTRY(unet_conv3(a.gda[l], Cl, nullptr, 0, la.wt, nullptr, nullptr, a.gcat[l], n, Hl, Wl,
                    Cl + Cn, 1, s));
    __bf16* gl = (l + 1 <= Lv - 2) ? a.gd[l + 1] : a.ge[Lv - 1];
    UTRY(unet_up_bwd(a.gcat[l], Cl, low, gl, n, Hl / 2, Wl / 2, Cn, s));
    if (l + 1 > Lv - 2) ge_top = gl;
  }
  // encoder, bottleneck first; ge[l]: gradient at the pre-activation of e_l
  (void)ge_top;
  for (int l = Lv - 1; l >= 0; --l) {
    const int Hl = H >> l, Wl = W >> l;
    const long Pl = (long)n * Hl * Wl;
    const int Cl = h->C(l);
    Layer& la = h->enc(l, 0);
    Layer& lb = h->enc(l, 1);
    UTRY(unet_conv3_dw(a.ge[l], Cl, a.ea[l], Cl, nullptr, 0, lb.gw, n, Hl, Wl, s));
    UTRY(unet_bias_grad(a.ge[l], lb.gb, Pl, Cl, s));
    UTRY(unet_conv3(a.ge[l], Cl, nullptr, 0, lb.wt, nullptr, a.ea[l], a.gea[l], n, Hl, Wl, Cl, 1,
                    s));
    if (l == 0) {
      UTRY(unet_conv3_in_dw(x, a.gea[0], la.gw, la.gb, n, H, W, Cl, s));
      break;
    }
    const int Cp = h->C(l - 1);
    UTRY(unet_conv3_dw(a.gea[l], Cl, a.p[l - 1], Cp, nullptr, 0, la.gw, n, Hl, Wl, s));
    UTRY(unet_bias_grad(a.gea[l], la.gb, Pl, Cl, s));
    UTRY(unet_conv3(a.gea[l], Cl, nullptr, 0, la.wt, nullptr, nullptr, a.gp[l - 1], n, Hl, Wl, Cp,
                    1, s));
    UTRY(unet_pool_bwd(a.e[l - 1], a.gp[l - 1], a.gcat[l - 1], Cp + Cl, a.ge[l - 1], n, H >> (l - 1),
                       W >> (l - 1), Cp, s));
  }
  return GLSU_OK;
}

// ================================================================ C ABI
static void unet_free(glsu_unet* h) {
  if (h->s) hipStreamSynchronize(h->s);
  acts_free(h->a);
  for (auto& l : h->L) {
    hipFree(l.w); hipFree(l.b); hipFree(l.gw); hipFree(l.gb);
    hipFree(l.mw); hipFree(l.vw); hipFree(l.mb); hipFree(l.vb);
    hipFree(l.wf); hipFree(l.wt);
  }
}

extern "C" int glsu_unet_create(glsu_unet** out, int levels, int base, unsigned long long seed,
                                void* stream) {
  if (!out || levels < 2 || levels > 6 || base < 32 || base % 32)
    return ufail(GLSU_ERR_ARG, "levels must be in 2..6 and base a multiple of 32");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return ufail(GLSU_ERR_DEVICE, "no HIP device");
  auto* h = new glsu_unet();
  h->levels = levels;
  h->base = base;
  h->s = (hipStream_t)stream;
  // layer list in the order of the reference definition
  auto add = [&](int cin, int cout, int k) {
    Layer l;
    l.cin = cin; l.cout = cout; l.k = k;
    h->L.push_back(l);
  };
  for (int l = 0; l < levels; ++l) {
    add(l == 0 ? 2 : h->C(l - 1), h->C(l), 3);
    add(h->C(l), h->C(l), 3);
  }
  for (int l = levels - 2; l >= 0; --l) {
    add(h->C(l) + h->C(l + 1), h->C(l), 3);
    add(h->C(l), h->C(l), 3);
  }
  add(h->C(0), 2, 1);
  std::mt19937_64 rng(seed);
  for (size_t i = 0; i < h->L.size(); ++i) {
    Layer& l = h->L[i];
    const long nw = l.nw();
    bool ok = hipMalloc((void**)&l.w, nw * 4) == hipSuccess &&
              hipMalloc((void**)&l.b, l.cout * 4) == hipSuccess &&
              hipMalloc((void**)&l.gw, nw * 4) == hipSuccess &&
              hipMalloc((void**)&l.gb, l.cout * 4) == hipSuccess &&
              hipMalloc((void**)&l.mw, nw * 4) == hipSuccess &&
              hipMalloc((void**)&l.vw, nw * 4) == hipSuccess &&
              hipMalloc((void**)&l.mb, l.cout * 4) == hipSuccess &&
              hipMalloc((void**)&l.vb, l.cout * 4) == hipSuccess;
    const bool mfma = l.k == 3 && l.cin % 32 == 0;
    if (ok && mfma)
      ok = hipMalloc((void**)&l.wf, nw * 2) == hipSuccess &&
           hipMalloc((void**)&l.wt, nw * 2) == hipSuccess;
    if (!ok) {
      unet_free(h);
      delete h;
      return ufail(GLSU_ERR_DEVICE, "weight allocation failed");
    }
    // He-normal weights, zero biases; the output layer starts at zero (identity net)
    std::vector<float> w(nw, 0.f), z(l.cout, 0.f);
    if (i + 1 < h->L.size()) {
      std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (l.k * l.k * l.cin)));
      for (auto& v : w) v = (float)nd(rng);
    }
    std::vector<float> zw(nw, 0.f);
    ok = hipMemcpy(l.w, w.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(l.b, z.data(), l.cout * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(l.mw, zw.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(l.vw, zw.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(l.mb, z.data(), l.cout * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(l.vb, z.data(), l.cout * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
      unet_free(h);
      delete h;
      return ufail(GLSU_ERR_DEVICE, "weight upload failed");
    }
  }
  if (int rc = refresh_bf16(h)) {
    unet_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return GLSU_OK;
}

extern "C" int glsu_unet_param_count(glsu_unet* h, long* count) {
  if (!h || !count) return ufail(GLSU_ERR_ARG, "null argument");
  long c = 0;
  for (auto& l : h->L) c += l.nw() + l.cout;
  *count = c;
  return GLSU_OK;
}

static int copy_params(glsu_unet* h, float* host, bool grads) {
  UTRY(hipStreamSynchronize(h->s));
  for (auto& l : h->L) {
    UTRY(hipMemcpy(host, grads ? l.gw : l.w, l.nw() * 4, hipMemcpyDeviceToHost));
    host += l.nw();
    UTRY(hipMemcpy(host, grads ? l.gb : l.b, l.cout * 4, hipMemcpyDeviceToHost));
    host += l.cout;
  }
  return GLSU_OK;
}

extern "C" int glsu_unet_get_params(glsu_unet* h, float* host) {
  if (!h || !host) return ufail(GLSU_ERR_ARG, "null argument");
  return copy_params(h, host, false);
}

extern "C" int glsu_unet_get_grads(glsu_unet* h, float* host) {
  if (!h || !host) return ufail(GLSU_ERR_ARG, "null argument");
  return copy_params(h, host, true);
}

extern "C" int glsu_unet_set_params(glsu_unet* h, const float* host) {
  if (!h || !host) return ufail(GLSU_ERR_ARG, "null argument");
  UTRY(hipStreamSynchronize(h->s));
  for (auto& l : h->L) {
    UTRY(hipMemcpy(l.w, host, l.nw() * 4, hipMemcpyHostToDevice));
    host += l.nw();
    UTRY(hipMemcpy(l.b, host, l.cout * 4, hipMemcpyHostToDevice));
    host += l.cout;
  }
  return refresh_bf16(h);
}

extern "C" int glsu_unet_forward(glsu_unet* h, const void* d_x, void* d_out, int n, int H,
                                 int W) {
  if (!h || !d_x || !d_out) return ufail(GLSU_ERR_ARG, "null argument");
  if (int rc = check_shape(h, n, H, W)) return rc;
  if (int rc = acts_alloc(h, n, H, W)) return rc;
  return forward(h, (const float2*)d_x, (float2*)d_out, n, H, W);
}

extern "C" int glsu_unet_loss_grad(glsu_unet* h, const void* d_x, const void* d_u, int n, int H,
                                   int W, double* loss) {
  if (!h || !d_x || !d_u) return ufail(GLSU_ERR_ARG, "null argument");
  if (int rc = check_shape(h, n, H, W)) return rc;
  if (int rc = acts_alloc(h, n, H, W)) return rc;
  if (int rc = forward(h, (const float2*)d_x, h->a.out, n, H, W)) return rc;
  if (int rc = backward(h, (const float2*)d_x, (const float2*)d_u, n, H, W)) return rc;
  if (loss) {
    float l = 0.f;
    UTRY(hipMemcpyAsync(&l, h->a.loss, 4, hipMemcpyDeviceToHost, h->s));
    UTRY(hipStreamSynchronize(h->s));
    *loss = l;
  }
  return GLSU_OK;
}

extern "C" int glsu_unet_adam_step(glsu_unet* h, float lr, float beta1, float beta2, float eps) {
  if (!h || !(lr > 0)) return ufail(GLSU_ERR_ARG, "invalid argument");
  ++h->step;
  for (auto& l : h->L) {
    UTRY(unet_adam(l.w, l.gw, l.mw, l.vw, l.nw(), lr, beta1, beta2, eps, h->step, h->s));
    UTRY(unet_adam(l.b, l.gb, l.mb, l.vb, l.cout, lr, beta1, beta2, eps, h->step, h->s));
  }
  return refresh_bf16(h);
}

extern "C" int glsu_unet_destroy(glsu_unet* h) {
  if (!h) return GLSU_OK;
  unet_free(h);
  delete h;
  return GLSU_OK;
}

namespace glsu {
// forward on the handle's activations, on stream s (used by the learned DPRA frame)
int unet_forward_on(glsu_unet* h, const float2* x, float2* out, int H, int W, hipStream_t s) {
  if (int rc = check_shape(h, 1, H, W)) return rc;
  if (int rc = acts_alloc(h, 1, H, W)) return rc;
  hipStream_t keep = h->s;
  h->s = s;
  int rc = forward(h, x, out, 1, H, W);
  h->s = keep;
  return rc;
}
}  // namespace glsu

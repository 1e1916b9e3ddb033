// unet_internal.h -- launchers of the U-Net kernels (unet_kernels.hip).
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

namespace glsu {

// 3x3 conv on BF16 MFMA over cat(X1, up2(X2)) (C1, C2 multiples of 32; X2 may
// be null with C2 = 0).  mode 0: Y = relu(conv + bias) with Wt [Cout][9][Cin];
// mode 1: Y = conv * (mask > 0) (mask may be null), no bias.  pool (mode 0,
// H and W even): also the 2 x 2 max-pool of Y, n x H/2 x W/2 x Cout, equal to
// unet_maxpool2(Y) (written from the conv's epilogue where the kernel allows).
hipError_t unet_conv3(const __bf16* X1, int C1, const __bf16* X2, int C2, const __bf16* Wt,
                      const float* bias, const __bf16* mask, __bf16* Y, int n, int H, int W,
                      int Cout, int mode, hipStream_t s, __bf16* pool = nullptr);
// dW [Cout][9][Cin] += weight gradient (FP32, atomically accumulated)
hipError_t unet_conv3_dw(const __bf16* dY, int Cout, const __bf16* X1, int C1, const __bf16* X2,
                         int C2, float* dW, int n, int H, int W, hipStream_t s);
hipError_t unet_conv3_in(const float2* x, const float* w, const float* b, __bf16* Y, int n,
                         int H, int W, int C0, hipStream_t s);
hipError_t unet_conv3_in_dw(const float2* x, const __bf16* dY, float* dW, float* db, int n,
                            int H, int W, int C0, hipStream_t s);
hipError_t unet_conv1_out(const __bf16* d, const float2* x, const float* w, const float* b,
                          float2* out, long P, int C0, hipStream_t s);
hipError_t unet_conv1_out_bwd(const __bf16* d, const float2* out, const float2* u,
                              const float* w, __bf16* dd, float* dW, float* db, float* loss,
                              long P, int C0, hipStream_t s);
hipError_t unet_maxpool2(const __bf16* x, __bf16* y, int n, int H, int W, int C, hipStream_t s);
hipError_t unet_pool_bwd(const __bf16* e, const __bf16* gp, const __bf16* gskip, int Cs,
                         __bf16* ge, int n, int H, int W, int C, hipStream_t s);
hipError_t unet_up_bwd(const __bf16* gcat, int C1, const __bf16* a, __bf16* ga, int n, int H2,
                       int W2, int C, hipStream_t s);
hipError_t unet_bias_grad(const __bf16* g, float* db, long P, int C, hipStream_t s);
hipError_t unet_adam(float* w, const float* g, float* m, float* v, long count, float lr,
                     float b1, float b2, float eps, int step, hipStream_t s);
hipError_t unet_weights_bf16(const float* w, __bf16* wf, __bf16* wt, int Cout, int Cin,
                             hipStream_t s);
hipError_t unet_fill(float* p, long count, float v, hipStream_t s);

}  // namespace glsu

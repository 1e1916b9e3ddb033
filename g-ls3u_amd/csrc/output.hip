// output.hip -- the output stage after the DPRA chain: cumulative phase
// change to 8-bit video frames, as the reference demo writes them
// (App_DPRA/dprawidget.cpp:541-549):
//   v = uchar((atan2(sin(p), cos(p)) + pi) / 2 / pi * 255)
// in FP32 with the same operation order; the conversion truncates.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/glsu.h"

namespace glsu {
void set_error(const char* msg);
}

namespace {

__device__ __forceinline__ uint8_t phase_byte(float p) {
  const float w = atan2f(sinf(p), cosf(p));
  const float v = (w + (float)M_PI) / 2.0f / (float)M_PI * 255.0f;
  return (uint8_t)v;
}

// Four pixels per thread (one 16-byte load, one 4-byte store) where aligned;
// the tail and unaligned buffers go one pixel at a time.
__global__ __launch_bounds__(256) void phase_to_u8_kernel(const float* __restrict__ phi,
                                                          uint8_t* __restrict__ out, size_t n,
                                                          int vec) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vec) {
    const size_t nq = n / 4;
    for (size_t q = i; q < nq; q += stride) {
      const float4 p = reinterpret_cast<const float4*>(phi)[q];
      const uint32_t w = (uint32_t)phase_byte(p.x) | ((uint32_t)phase_byte(p.y) << 8) |
                         ((uint32_t)phase_byte(p.z) << 16) | ((uint32_t)phase_byte(p.w) << 24);
      reinterpret_cast<uint32_t*>(out)[q] = w;
    }
    for (size_t k = nq * 4 + i; k < n; k += stride) out[k] = phase_byte(phi[k]);
  } else {
    for (size_t k = i; k < n; k += stride) out[k] = phase_byte(phi[k]);
  }
}

}  // namespace

extern "C" int glsu_phase_to_u8(const float* d_phi, uint8_t* d_out, size_t n, void* hip_stream) {
  if (n == 0) return GLSU_OK;
  if (!d_phi || !d_out) {
    glsu::set_error("null argument");
    return GLSU_ERR_ARG;
  }
  const int vec = ((uintptr_t)d_phi % 16 == 0) && ((uintptr_t)d_out % 4 == 0);
  size_t blocks = (n / 4 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(phase_to_u8_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)hip_stream, d_phi, d_out, n, vec);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof buf, "phase_to_u8 launch failed: %s", hipGetErrorString(e));
    glsu::set_error(buf);
    return GLSU_ERR_DEVICE;
  }
  return GLSU_OK;
}

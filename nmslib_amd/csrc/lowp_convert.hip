// Reduced-precision copies of the level-0 vectors, built on the device from
// the FP32 vectors already resident in HBM.
//   BF16: round to nearest even of the FP32 value (NaN-free inputs).
//   FP8 : e4m3 with one FP32 scale per vector, scale = max|x| / 240 so every
//         element lands inside the finite range of the e4m3 encodings the
//         hardware converter uses; a zero vector keeps scale 1.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hnsw_device.h"

namespace nmsamd {

__global__ void to_bf16_kernel(const float* __restrict__ vec, uint16_t* __restrict__ out,
                               int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t b = __float_as_uint(vec[i]);
  const uint32_t r = b + 0x7fffu + ((b >> 16) & 1u);
  out[i] = (uint16_t)(r >> 16);
}

// one wave per vector
__global__ void to_fp8_kernel(const float* __restrict__ vec, uint8_t* __restrict__ out,
                              float* __restrict__ scale, int64_t n, int dim, int dstride) {
  const int64_t v = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= n) return;
  const float* x = vec + v * dstride;
  float amax = 0.f;
  for (int i = lane; i < dim; i += 64) amax = fmaxf(amax, fabsf(x[i]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
  const float s = amax > 0.f ? amax / 240.0f : 1.0f;
  const float inv = 1.0f / s;
  for (int i = 2 * lane; i < dstride; i += 128) {
    const float a = i < dim ? x[i] * inv : 0.f;
    const float b = i + 1 < dim ? x[i + 1] * inv : 0.f;
    const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    out[v * dstride + i] = (uint8_t)(packed & 0xff);
    if (i + 1 < dstride) out[v * dstride + i + 1] = (uint8_t)((packed >> 8) & 0xff);
  }
  if (lane == 0) scale[v] = s;
}

hipError_t launch_convert_lowp(const float* vec, uint16_t* vec16, uint8_t* vec8, float* scale8,
                               int64_t n, int dim, int dstride, int storage, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (storage == kBF16) {
    const int64_t total = n * dstride;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       vec, vec16, total);
  } else {
    hipLaunchKernelGGL(to_fp8_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, stream, vec,
                       vec8, scale8, n, dim, dstride);
  }
  return hipGetLastError();
}

}  // namespace nmsamd

// Launch entries used by the host library: kernel selection by vector
// storage, dimension, distance function and register width, plus a small
// fill kernel (used instead of hipMemsetAsync so a captured launch sequence
// holds only kernel nodes).
#include <hip/hip_runtime.h>

#include "hnsw_device.h"

namespace nmsamd {

using KernelFn = void (*)(SearchArgs, int);
KernelFn pick_kernel_f32(int dim, int func, int mode);
KernelFn pick_kernel_bf16(int dim, int func, int mode);
KernelFn pick_kernel_fp8(int dim, int func, int mode);

hipError_t launch_search(const SearchArgs& a, int blocks, int waves_per_block, int lds_wave_bytes,
                         int mode, hipStream_t stream) {
  KernelFn fn = nullptr;
  if (a.storage == kF32) fn = pick_kernel_f32(a.ix.dim, a.ix.func, mode);
  else if (a.storage == kBF16) fn = pick_kernel_bf16(a.ix.dim, a.ix.func, mode);
  else fn = pick_kernel_fp8(a.ix.dim, a.ix.func, mode);
  if (!fn) return hipErrorInvalidValue;
  if (waves_per_block < 1 || waves_per_block > kWavesPerBlock) return hipErrorInvalidValue;
  // kModeBeamPair: one query per block, served by two waves on one LDS area
  const bool pair = mode == kModeBeamPair;
  if (pair && waves_per_block != 1) return hipErrorInvalidValue;
  const size_t lds = (size_t)lds_wave_bytes * waves_per_block;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWave * (pair ? 2 : waves_per_block)), lds, stream, a,
                     lds_wave_bytes);
  return hipGetLastError();
}

__global__ void fill_i32_kernel(int* p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int* p, int v, int n, hipStream_t stream) {
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

}  // namespace nmsamd

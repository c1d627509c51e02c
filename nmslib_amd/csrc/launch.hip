// Launch entries used by the host library: kernel selection by vector
// storage, dimension, distance function and register width, plus a small
// fill kernel (used instead of hipMemsetAsync so a captured launch sequence
// holds only kernel nodes).
#include <hip/hip_runtime.h>

#include "hnsw_device.h"

namespace nmsamd {

using KernelFn = void (*)(SearchArgs, int);
// the slices of search_inst.hip: pick_kernel_st<storage>_m<mode>
#define NMS_DECL(S)                                    \
  KernelFn pick_kernel_st##S##_m0(int dim, int func);  \
  KernelFn pick_kernel_st##S##_m1(int dim, int func);  \
  KernelFn pick_kernel_st##S##_m2(int dim, int func);  \
  KernelFn pick_kernel_st##S##_m3(int dim, int func);  \
  KernelFn pick_kernel_st##S##_m4(int dim, int func);  \
  KernelFn pick_kernel_st##S##_m5(int dim, int func);
NMS_DECL(0)
NMS_DECL(1)
NMS_DECL(2)
#undef NMS_DECL

static KernelFn pick_kernel(int storage, int dim, int func, int mode) {
  using Pick = KernelFn (*)(int, int);
  static const Pick table[3][6] = {
      {pick_kernel_st0_m0, pick_kernel_st0_m1, pick_kernel_st0_m2, pick_kernel_st0_m3,
       pick_kernel_st0_m4, pick_kernel_st0_m5},
      {pick_kernel_st1_m0, pick_kernel_st1_m1, pick_kernel_st1_m2, pick_kernel_st1_m3,
       pick_kernel_st1_m4, pick_kernel_st1_m5},
      {pick_kernel_st2_m0, pick_kernel_st2_m1, pick_kernel_st2_m2, pick_kernel_st2_m3,
       pick_kernel_st2_m4, pick_kernel_st2_m5}};
  if (storage < 0 || storage > 2 || mode < 0 || mode > 5) return nullptr;
  return table[storage][mode](dim, func);
}

hipError_t launch_search(const SearchArgs& a, int blocks, int waves_per_block, int lds_wave_bytes,
                         int mode, hipStream_t stream) {
  KernelFn fn = pick_kernel(a.storage, a.ix.dim, a.ix.func, mode);
  if (!fn) return hipErrorInvalidValue;
  if (waves_per_block < 1 || waves_per_block > kWavesPerBlock) return hipErrorInvalidValue;
  // kModeBeamPair / kModeBeamCoop: one query per block, served by two /
  // kCoopWaves waves on one LDS area
  const bool pair = mode == kModeBeamPair, coop = mode == kModeBeamCoop || mode == kModeBeamCoop2;
  if ((pair || coop) && waves_per_block != 1) return hipErrorInvalidValue;
  const size_t lds = (size_t)lds_wave_bytes * waves_per_block;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int threads = kWave * (coop ? kCoopWaves : (pair ? 2 : waves_per_block));
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), lds, stream, a, lds_wave_bytes);
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

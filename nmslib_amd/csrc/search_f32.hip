// Instantiations of the search kernels for f32 vector storage.
#include "hnsw_search_impl.h"

namespace nmsamd {

KernelFn pick_kernel_f32(int dim, int func, int mode) {
#define NMS_CASE(D, W)                                                 \
  switch (func) {                                                      \
    case kL2Sqr16: return kernel_of<D, kL2Sqr16, kF32, W>();      \
    case kL2Sqr: return kernel_of<D, kL2Sqr, kF32, W>();          \
    case kNormCos: return kernel_of<D, kNormCos, kF32, W>();      \
    case kNegDot: return kernel_of<D, kNegDot, kF32, W>();        \
    case kL1: return kernel_of<0, kL1, kF32, W>();                 \
    case kLInf: return kernel_of<0, kLInf, kF32, W>();             \
    case kSpaceL2: return kernel_of<0, kSpaceL2, kF32, W>();       \
    case kSpaceCos: return kernel_of<0, kSpaceCos, kF32, W>();     \
    case kSpaceNegDot: return kernel_of<0, kSpaceNegDot, kF32, W>(); \
    default: return nullptr;                                           \
  }
#define NMS_DIMS(W)                \
  switch (dim) {                   \
    case 96: NMS_CASE(96, W)       \
    case 100: NMS_CASE(100, W)     \
    case 128: NMS_CASE(128, W)     \
    default: NMS_CASE(0, W)        \
  }
  if (mode == kModeBeamPair) {
    NMS_DIMS(kModeBeamPair)
  } else if (mode == kModeBeamWide) {
    NMS_DIMS(kModeBeamWide)
  } else if (mode == kModeBeam) {
    NMS_DIMS(kModeBeam)
  } else {
    NMS_DIMS(kModeAll)
  }
#undef NMS_DIMS
#undef NMS_CASE
}

}  // namespace nmsamd

// Instantiations of the search kernels for bf16 vector storage.
#include "hnsw_search_impl.h"

namespace nmsamd {

KernelFn pick_kernel_bf16(int dim, int func, int mode) {
#define NMS_CASE(D, W)                                                 \
  switch (func) {                                                      \
    case kL2Sqr16: return kernel_of<D, kL2Sqr16, kBF16, W>();      \
    case kL2Sqr: return kernel_of<D, kL2Sqr, kBF16, W>();          \
    case kNormCos: return kernel_of<D, kNormCos, kBF16, W>();      \
    case kNegDot: return kernel_of<D, kNegDot, kBF16, W>();        \
    case kL1: return kernel_of<0, kL1, kBF16, W>();                 \
    case kLInf: return kernel_of<0, kLInf, kBF16, W>();             \
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

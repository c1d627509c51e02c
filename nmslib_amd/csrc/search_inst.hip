// One slice of the search kernel instantiations: vector storage NMS_ST
// (0 FP32, 1 BF16, 2 FP8) x kernel mode NMS_MODE (0 every mode, 1 beam,
// 2 beam wide, 3 beam two waves, 4 beam cooperative, 5 beam cooperative at two
// blocks per CU), over the specialised dimensions {96, 100,
// 128, any} and the distance functions.  The build compiles this file once per
// slice so the eighteen slices compile in parallel; launch.hip picks the slice.
#include "hnsw_search_impl.h"

#if !defined(NMS_ST) || !defined(NMS_MODE)
#error "NMS_ST and NMS_MODE select the slice"
#endif

#define NMS_CAT2(a, b) a##b
#define NMS_CAT(a, b) NMS_CAT2(a, b)
#define NMS_PICK NMS_CAT(NMS_CAT(NMS_CAT(pick_kernel_st, NMS_ST), _m), NMS_MODE)

namespace nmsamd {

KernelFn NMS_PICK(int dim, int func) {
  constexpr int ST = NMS_ST;
  constexpr int W = NMS_MODE;
#define NMS_CASE(D)                                                        \
  switch (func) {                                                          \
    case kL2Sqr16: return kernel_of<D, kL2Sqr16, ST, W>();                 \
    case kL2Sqr: return kernel_of<D, kL2Sqr, ST, W>();                     \
    case kNormCos: return kernel_of<D, kNormCos, ST, W>();                 \
    case kNegDot: return kernel_of<D, kNegDot, ST, W>();                   \
    case kL1: return kernel_of<0, kL1, ST, W>();                           \
    case kLInf: return kernel_of<0, kLInf, ST, W>();                       \
    default: break;                                                        \
  }                                                                        \
  /* the spaces' own distances of a regular index: FP32 only */            \
  if constexpr (ST == kF32) {                                              \
    switch (func) {                                                        \
      case kSpaceL2: return kernel_of<0, kSpaceL2, ST, W>();               \
      case kSpaceCos: return kernel_of<0, kSpaceCos, ST, W>();             \
      case kSpaceNegDot: return kernel_of<0, kSpaceNegDot, ST, W>();       \
      default: break;                                                      \
    }                                                                      \
  }                                                                        \
  return nullptr;
  switch (dim) {
    case 96: { NMS_CASE(96) }
    case 100: { NMS_CASE(100) }
    case 128: { NMS_CASE(128) }
    default: { NMS_CASE(0) }
  }
#undef NMS_CASE
}

}  // namespace nmsamd

// Host-side HNSW index construction: the graph nmslib's Hnsw::CreateIndex
// builds (similarity_search/src/method/hnsw.cc:183-470), laid out as the
// optimized index that LoadOptimizedIndex reads (hnsw.cc:748-806,
// include/method/hnsw.h:328-371).
#ifndef NMSLIB_AMD_HNSW_BUILD_H
#define NMSLIB_AMD_HNSW_BUILD_H
#include <cstddef>
#include <cstdint>
#include <vector>

namespace nmsamd {

// Index-time distance of each dense space (the spaces' HiddenDistance).
enum BuildSpace : int {
  kBuildL2 = 1,      // SpaceLp(2): L2NormSIMD, the square root of the squared distance
  kBuildCosine = 3,  // SpaceCosineSimilarity: max(0, 1 - NormScalarProductSIMD)
  kBuildNegDot = 4,  // SpaceNegativeScalarProduct: -ScalarProductSIMD
  kBuildL1 = 5,      // SpaceLp(1): L1NormSIMD
  kBuildLInf = 6,    // SpaceLp(-1): LInfNormSIMD
};

// The CreateIndex parameters (hnsw.cc:187-205) after defaults are applied.
struct BuildParams {
  size_t M = 16;
  size_t maxM = 16;       // upper-level list capacity (default M)
  size_t maxM0 = 32;      // level-0 list capacity (default 2M)
  size_t efConstruction = 200;
  size_t threads = 1;     // indexThreadQty
  double mult = 0;        // level multiplier (default 1/ln M)
  int delaunay = 2;       // neighbour selection: 0 closest, heuristic 1, 2 or 3
  int post = 0;           // 1, 2: merge with a graph built in reverse order
  bool progress = false;  // print_progress: report inserted points to stderr
  bool optimized = true;  // false (skip_optimized_index=1): keep the regular form, whose
                          // vectors are not normalised and whose distances are the space's
};

// An optimized index image: the file's header fields, its level-0 block
// (N elements of `elem` bytes) and the per-node upper-level lists.
struct BuiltIndex {
  uint32_t n = 0;
  uint64_t elem = 0, off_l0 = 0, maxM = 0, maxM0 = 0;
  int32_t maxlevel = 0, func = 0;
  uint32_t enter = 0;
  std::vector<char> block;
  std::vector<uint32_t> sizemass;   // bytes of upper-level lists per node
  std::vector<int32_t> upper;       // concatenated [count, ids[maxM]] per level >= 1
  std::vector<int32_t> upper_off;   // start of a node's lists in `upper`, -1 when none
};

// Space::IndexTimeDistance(a, b) of two dense float vectors of `dim` floats
// (the space's own distance, in the reference's SSE order).
float space_distance(int space, const float* a, const float* b, size_t dim);

// Builds from `objs`, the data points as the reference stores them (Object
// buffers: i32 id, i32 label, u64 datalength, f32 data[dim]), in insertion
// order.  Throws std::invalid_argument / std::runtime_error on bad input.
void build_hnsw(int space, const std::vector<std::vector<char>>& objs, const BuildParams& p,
                BuiltIndex& out);

}  // namespace nmsamd
#endif

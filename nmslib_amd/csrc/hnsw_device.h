// Device-resident HNSW index and launch arguments shared by the host library
// and the HIP kernels.
//
// HBM layout of a loaded optimized index (one copy per device, read-only
// while searching):
//   vec     [n][dstride] f32   level-0 vectors as stored in the index file
//                              (cosine ones already normalised there); rows
//                              are padded with zeros to a multiple of 4 floats
//                              so every row starts on a 16-byte boundary
//   vec16   [n][dstride] bf16  optional reduced-precision copy
//   vec8    [n][dstride] fp8   optional e4m3 copy, with scale8[n] f32
//   l0      [n][l0stride] i32  level-0 adjacency, the list's ids in stored
//                              order followed by -1 up to the padded stride
//                              (l0stride = maxM0 rounded up to 32 -> 128 B rows)
//   up_off  [n] i32            offset of a node's upper-level lists in `up`,
//                              -1 for level-0-only nodes
//   up      flat i32           per node, per level L>=1: [count, ids[maxM]]
//                              (the file's own per-level blocks)
//   ext_id  [n] i32            object id stored in each element (results
//                              report these, hnsw_distfunc_opt.cc:280)
// Per search slot (a resident wavefront):
//   visited [slots][vstride] u8  epoch marks, as the reference's VisitedList
//   epoch   [slots] u8           current epoch of each slot
#ifndef NMSLIB_AMD_HNSW_DEVICE_H
#define NMSLIB_AMD_HNSW_DEVICE_H
#include <stdint.h>

namespace nmsamd {

// distance function codes of the optimized index (include/method/hnsw.h:50-58),
// then the spaces' own distances that a regular (non-optimized) index is
// searched with (hnsw.cc:1076-1300 through Query::DistanceObjLeft: the data
// point is the left operand): L2NormSIMD (the square root of the squared
// distance), CosineSimilarity of the raw vectors and -ScalarProductSIMD, all in
// the four-lane SSE order.  l1 and linf use codes 5 and 6 in both formats.
enum DistFunc : int {
  kL2Sqr16 = 1, kL2Sqr = 2, kNormCos = 3, kNegDot = 4, kL1 = 5, kLInf = 6,
  kSpaceL2 = 7, kSpaceCos = 8, kSpaceNegDot = 9
};
// vector storage used while traversing
enum Storage : int { kF32 = 0, kBF16 = 1, kFP8 = 2 };

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kMaxM0 = 64;          // widest level-0 list one expansion handles
constexpr uint32_t kUsedBit = 0x80000000u;

struct DevIndex {
  const float* vec;
  const uint16_t* vec16;
  const uint8_t* vec8;
  const float* scale8;
  const int32_t* l0;
  const int32_t* up_off;
  const int32_t* up;
  const int32_t* ext_id;
  uint8_t* visited;
  uint8_t* epoch;
  int64_t n;
  int64_t vstride;
  int dim, dstride, l0stride, maxM, maxM0, maxlevel, enter, func;
  int has_dups;
};

struct SearchStats {
  unsigned long long expansions;
  unsigned long long dist_level0;
  unsigned long long dist_upper;
  unsigned long long list_ints_read;
  unsigned long long tie_paths;
  // beam mode: 100 MHz ticks summed over queries per phase of a step
  // (pick, adjacency+visited, distances, filter, sort+prediction, rank search,
  // rebuild, advance)
  unsigned long long phase_ticks[8];
  // the BF16 screen (cooperative FP32 l2 search): candidates screened, and
  // the ones that passed to their FP32 distance
  unsigned long long screened;
  unsigned long long screen_pass;
};

struct SearchArgs {
  DevIndex ix;
  const float* queries;      // [nq][dim]
  int64_t nq;
  const int32_t* k_of;       // optional per-query k
  const int32_t* ef_of;      // optional per-query ef
  int k, ef;                 // used when the per-query arrays are null
  int cap;                   // LDS capacity of the sorted array (>= max(ef,k) of every query)
  int out_stride;            // row stride of out_ids / out_dists
  int32_t* out_ids;
  float* out_dists;
  int32_t* out_counts;
  int32_t* out_internal;     // optional [nq][out_stride] internal node ids
  uint32_t* q_ticks;         // optional [nq] per-query search time, 100 MHz ticks
  int* work;                 // work counter, zero before the launch
  SearchStats* stats;        // optional
  int storage;               // Storage
  int rerank;                // > 0: exact FP32 rerank of the best rerank*k candidates
  int beam;                  // > 0: beam mode, expansions per step (extension)
  int ent;                   // per-step neighbour entries the LDS scratch holds
                             // (a multiple of 64, at least 128)
  int64_t slot_base;        // first visited-mark slot of this launch
  const int32_t* order;      // optional: the order in which waves take queries
  const int64_t* nq_dev;     // optional: the query count read on the device
                             // (a replayed graph keeps its kernel arguments)
  int algo;                  // AlgoMode
  float* old_heap_key;       // SearchOld candidate heap, old_heap_cap per slot
  int32_t* old_heap_id;
  int64_t old_heap_cap;
  int* err;                  // bit 1: a SearchOld candidate heap overflowed; bit 2: the
                             // search state held a node id outside the index or expanded an item twice;
                             // bit 4: the screen's bound failed (verify mode)
  int defer;                 // two-wave beam queries: the first wave rebuilds the sorted
                             // array while the second computes the next step's distances
                             // (needs the second items buffer after the command words)
  int vbits;                 // visited marks as bits (cleared per query), else epoch bytes
  int screen;                // cooperative FP32 l2 beam search: 1 = screen candidates by a
                             // lower bound from their BF16 copy, 2 = and verify the bound
                             // against every FP32 distance (err bit 4 if it failed)
};

// algoType (hnsw.cc:491-501); hybrid picks SearchOld for ef >= 1000 per query
enum AlgoMode : int { kAlgoHybrid = 0, kAlgoOld = 1, kAlgoV1Merge = 2 };

// Kernel variants: every search mode, or beam mode only (lower register
// pressure), the latter also with all of a step's vector loads in flight, or
// with two waves per query (a block of two waves: the second computes each
// step's distances, and the first rebuilds the sorted array meanwhile or, where
// LDS lacks room for a second items buffer, each wave rebuilds half of it).
enum KernelMode : int {
  kModeAll = 0, kModeBeam = 1, kModeBeamWide = 2, kModeBeamPair = 3,
  kModeBeamCoop = 4,  // one query per block of kCoopWaves waves (beam up to 64)
  kModeBeamCoop2 = 5  // the same, two blocks per CU (registers of four waves per SIMD)
};
constexpr int kCoopWaves = 8;

}  // namespace nmsamd
#endif

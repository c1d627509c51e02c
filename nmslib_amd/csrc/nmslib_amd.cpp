// Host side of the MI355X HNSW query path: the IndexWrapper-like handle behind
// include/nmslib_amd.h.
//
// Reference counterparts (similarity_search/ and python_bindings/ of nmslib):
//   HnswIndex::load            Hnsw::LoadIndex / LoadOptimizedIndex hnsw.cc:993-1074
//   HnswIndex::save            Hnsw::SaveIndex / SaveOptimizedIndex hnsw.cc:748-806
//   HnswIndex::set_query_params Hnsw::SetQueryTimeParams hnsw.cc:472-507
//   HnswIndex::search          IndexWrapper::knnQueryBatch nmslib.cc:142-184
//                              + Hnsw::Search dispatch hnsw.cc:717-746
//   init                       nmslib.init nmslib.cc:470-510, spacefactory.h:46-54
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nmslib_amd.h"
#include "hnsw_build.h"
#include "hnsw_device.h"

namespace nmsamd {
hipError_t launch_search(const SearchArgs& a, int blocks, int waves_per_block, int lds_wave_bytes,
                         int mode, hipStream_t stream);
hipError_t launch_fill_i32(int* p, int v, int n, hipStream_t stream);
hipError_t launch_convert_lowp(const float* vec, uint16_t* vec16, uint8_t* vec8, float* scale8,
                               int64_t n, int dim, int dstride, int storage, hipStream_t stream);
}  // namespace nmsamd

namespace {

using namespace nmsamd;

thread_local std::string g_err;

struct Status : std::runtime_error {
  int code;
  Status(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Status(code, msg); }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    fail(NMSAMD_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    ensure(count);
    if (count)
      hip_check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
};

enum Algo { kHybrid = 0, kOld = 1, kV1Merge = 2 };

constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950 LDS per CU (and per workgroup)

std::string to_lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

template <class T>
void rd(std::istream& in, T& v) { in.read(reinterpret_cast<char*>(&v), sizeof(T)); }
template <class T>
void wr(std::ostream& out, const T& v) { out.write(reinterpret_cast<const char*>(&v), sizeof(T)); }

}  // namespace

struct nmsamd_index {
  std::string space;
  int expected_func = 0;  // 0: any l2 variant
  // query-time parameters (Hnsw defaults after LoadIndex, hnsw.cc:483)
  size_t ef = 20;
  Algo algo = kHybrid;
  int storage = kF32;
  int rerank_mult = 0;
  int beam = 0;  // 0: exact V1Merge; > 0: beam mode width
  int beam_waves = 0;  // waves per beam query: 0 = by LDS footprint, 1 or 2
  int beam_gather = 0;  // one-wave beam kernel: 0 = by LDS footprint, 1 = plain, 2 = wide
  int visited_marks = 0;  // 0 = by size (use_bit_marks), 1 = epoch bytes, 2 = bits
  int screen = 0;  // cooperative FP32 l2 beam search: 1 = BF16 screen, 2 = and verify
  bool slots_bits = false;  // the last launch over d_visited left bit marks in it
  bool collect = false;
  // file header (hnsw.cc:779-788)
  bool loaded = false;
  bool on_device = false;  // uploaded on the first search when no device was visible before
  int servers = 0;         // open serving handles: they hold this index's device pointers
  // regular (non-optimized) index (hnsw.cc:810-841, 943-991): searched with the
  // space's own distance; saved back in its own format
  bool regular = false;
  uint64_t reg_M = 0;          // M_ of its header
  int32_t file_maxlevel = 0;   // maxlevel_ of its header (the search starts at the entry's level)
  uint32_t n = 0;
  uint64_t elem = 0, off_l0 = 0, off_data = 0, maxM = 0, maxM0 = 0, method = 0;
  int32_t maxlevel = 0, func = 0;
  uint32_t enter = 0;
  int dim = 0, dstride = 0, l0stride = 0;
  // host copies kept for saveIndex (the reference keeps the same block in RAM)
  std::vector<char> block;
  std::vector<uint32_t> sizemass;
  std::vector<int32_t> upper;       // concatenated upper-level blocks
  std::vector<int32_t> upper_off;   // per node, -1 when none
  std::vector<std::vector<char>> data_objs;  // load_data objects (.dat)
  bool has_dups = false;
  // device copies
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf<float> d_vec;
  DevBuf<uint16_t> d_vec16;
  DevBuf<uint8_t> d_vec8;
  DevBuf<float> d_scale8;
  DevBuf<int32_t> d_l0, d_upoff, d_up, d_ext;
  DevBuf<uint8_t> d_visited, d_epoch;
  int64_t vstride = 0;
  int64_t vstride_server = 0;
  int slots = 0;
  // per-call buffers
  DevBuf<float> d_q, d_dists;
  DevBuf<int32_t> d_ids, d_counts, d_k, d_ef, d_internal, d_work;
  DevBuf<SearchStats> d_stats;
  DevBuf<uint32_t> d_qticks;
  DevBuf<float> d_oldk;
  DevBuf<int32_t> d_oldi;
  DevBuf<int32_t> d_order;  // uneven batches: query order
  const int32_t* order_dev = nullptr;
  // brute force: |x|^2, padded queries, per-query screen error bounds, partial
  // lists, screened candidates, listed queries (+ their count)
  DevBuf<float> d_xnorm, d_bq, d_beps, d_bpd, d_bcd;
  DevBuf<int32_t> d_bpi, d_bci, d_bflag;
  bool vec_bounds_known = false;
  double vec_norm_max = 0, vec_int_absmax = -1;
  int64_t last_bf_rescanned = 0;
  bool pending = false;
  hipStream_t last_stream = nullptr;  // stream of the last search launch (ev1 follows it)
  int* last_err = nullptr;            // error word of the last search launch
  bool last_collect = false;          // the last search launch collected counters
  int64_t last_nq = 0;
  SearchStats last_stats{};
  int64_t last_launch[5] = {0, 0, 0, 0, 0};  // kernel mode, blocks, waves/block, LDS/wave, entries
  double last_ms = 0;
  std::mutex mu;  // one search at a time per handle
  int cu_count = 0;

  ~nmsamd_index() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void ensure_device(nmsamd_index* h) {
  if (h->stream) return;
  int count = 0;
  hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (count <= 0) fail(NMSAMD_EDEVICE, "no HIP device is visible");
  hip_check(hipGetDevice(&h->device), "hipGetDevice");
  hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreate(&h->ev0), "hipEventCreate");
  hip_check(hipEventCreate(&h->ev1), "hipEventCreate");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, h->device), "hipGetDeviceProperties");
  h->cu_count = prop.multiProcessorCount;
}

// Space code -> distance code of a regular index (the space's own function).
int regular_func(int expected_func) {
  switch (expected_func) {
    case 1: return kSpaceL2;
    case kNormCos: return kSpaceCos;
    case kNegDot: return kSpaceNegDot;
    default: return expected_func;  // l1, linf: the same functions as the optimized codes
  }
}

// Lays a graph given as per-node lists out like an optimized index image over
// the handle's data points (vectors as given: the regular search does not
// normalise).  lists[i][L] = node i's list at level L.
void image_from_lists(nmsamd_index* h, const std::vector<std::vector<std::vector<int32_t>>>& lists) {
  const size_t n = lists.size();
  if (n != h->data_objs.size())
    fail(NMSAMD_ERUNTIME, "The number of stored elements " + std::to_string(n) +
                              " doesn't match the number of data points " +
                              std::to_string(h->data_objs.size()) + "! Did you forget to re-load data?");
  uint64_t len = 0;
  for (size_t i = 0; i < n; ++i) {
    const auto& o = h->data_objs[i];
    uint64_t dl = 0;
    if (o.size() >= 16) std::memcpy(&dl, o.data() + 8, 8);
    if (o.size() < 16 || dl + 16 != o.size() || dl == 0 || dl % 4 != 0 || (i && dl != len))
      fail(NMSAMD_ERUNTIME, "data point " + std::to_string(i) + " is not a dense float vector "
                            "of the index's dimension");
    len = dl;
  }
  h->dim = (int)(len / 4);
  h->off_l0 = 16 + len;
  h->off_data = 0;
  h->elem = h->off_l0 + 4 * (h->maxM0 + 1);
  h->block.assign(n * h->elem, 1);
  h->sizemass.assign(n, 0);
  h->upper.clear();
  h->upper_off.assign(n, -1);
  for (size_t i = 0; i < n; ++i) {
    char* e = h->block.data() + i * h->elem;
    std::memcpy(e, h->data_objs[i].data(), h->off_l0);
    const auto& l0 = lists[i][0];
    if (l0.size() > h->maxM0) fail(NMSAMD_ERUNTIME, "level-0 list longer than maxM0");
    const int32_t cnt = (int32_t)l0.size();
    std::memcpy(e + h->off_l0, &cnt, 4);
    if (cnt) std::memcpy(e + h->off_l0 + 4, l0.data(), 4 * l0.size());
    const size_t levels = lists[i].size() - 1;
    if (levels) {
      h->sizemass[i] = (uint32_t)(levels * (h->maxM + 1) * 4);
      h->upper_off[i] = (int32_t)h->upper.size();
      for (size_t L = 1; L <= levels; ++L) {
        const auto& lst = lists[i][L];
        if (lst.size() > h->maxM) fail(NMSAMD_ERUNTIME, "upper-level list longer than maxM");
        h->upper.push_back((int32_t)lst.size());
        h->upper.insert(h->upper.end(), lst.begin(), lst.end());
        h->upper.resize(h->upper.size() + (h->maxM - lst.size()), 0);
      }
    }
  }
  h->n = (uint32_t)n;
  h->func = regular_func(h->expected_func);
  h->method = 0;
  h->regular = true;
  // baseSearchAlgorithm*: the descent starts at the entry point's own level
  h->maxlevel = n ? (int32_t)(lists[h->enter].size() - 1) : 0;
}

// LoadRegularIndexBin (hnsw.cc:943-991), after the format flag: the header,
// then per node its level and, for levels 0..level, [count, ids].  The index
// holds no vectors: the handle's data points must be loaded or added first.
void read_regular_body(nmsamd_index* h, std::istream& in) {
  uint32_t n = 0, enter = 0;
  int32_t maxlevel = 0;
  uint64_t M = 0, maxM = 0, maxM0 = 0;
  rd(in, n);
  rd(in, maxlevel);
  rd(in, enter);
  rd(in, M);
  rd(in, maxM);
  rd(in, maxM0);
  if (n != h->data_objs.size())
    fail(NMSAMD_ERUNTIME, "The number of stored elements " + std::to_string(n) +
                              " doesn't match the number of data points " +
                              std::to_string(h->data_objs.size()) + "! Did you forget to re-load data?");
  if (maxM0 > (uint64_t)kMaxM0 || maxM > (uint64_t)kMaxM0)
    fail(NMSAMD_ERUNTIME, "maxM0 " + std::to_string(maxM0) + " exceeds the device limit of " +
                              std::to_string(kMaxM0));
  if (n && enter >= n) fail(NMSAMD_ERUNTIME, "entry point out of range");
  std::vector<std::vector<std::vector<int32_t>>> lists(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t level = 0;
    rd(in, level);
    if (level > 64) fail(NMSAMD_ERUNTIME, "corrupt node level");
    lists[i].resize(level + 1);
    for (uint32_t L = 0; L <= level; ++L) {
      uint32_t qty = 0;
      rd(in, qty);
      if (qty > (L ? maxM : maxM0)) fail(NMSAMD_ERUNTIME, "corrupt friend list size");
      auto& f = lists[i][L];
      f.resize(qty);
      if (qty) in.read(reinterpret_cast<char*>(f.data()), 4 * (size_t)qty);
      for (int32_t v : f)
        if (v < 0 || (uint32_t)v >= n)
          fail(NMSAMD_ERUNTIME, "Invalid friendId = " + std::to_string(v) +
                                    " for node id: " + std::to_string(i));
    }
  }
  h->maxM = maxM;
  h->maxM0 = maxM0;
  h->enter = enter;
  h->reg_M = M;
  h->file_maxlevel = maxlevel;
  image_from_lists(h, lists);
}

// SaveRegularIndexBin (hnsw.cc:810-841) from the handle's image.
void write_regular(const nmsamd_index* h, std::ostream& out) {
  wr(out, (uint32_t)0);
  wr(out, h->n);
  wr(out, h->file_maxlevel);
  wr(out, h->enter);
  wr(out, h->reg_M);
  wr(out, h->maxM);
  wr(out, h->maxM0);
  for (uint32_t i = 0; i < h->n; ++i) {
    const char* e = h->block.data() + (size_t)i * h->elem;
    const uint32_t levels = h->sizemass[i] / (uint32_t)(4 * (h->maxM + 1));
    wr(out, levels);
    int32_t cnt;
    std::memcpy(&cnt, e + h->off_l0, 4);
    wr(out, (uint32_t)cnt);
    out.write(e + h->off_l0 + 4, 4 * (size_t)cnt);
    for (uint32_t L = 1; L <= levels; ++L) {
      const int32_t* lst = h->upper.data() + h->upper_off[i] + (h->maxM + 1) * (L - 1);
      wr(out, (uint32_t)lst[0]);
      out.write(reinterpret_cast<const char*>(lst + 1), 4 * (size_t)lst[0]);
    }
  }
}


// Reads an index file (LoadIndex, hnsw.cc:993-1074: the optimized format, or
// the regular one over the handle's data points) into the handle's host copy.
void read_index_file(nmsamd_index* h, const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) fail(NMSAMD_EIO, "Cannot open file '" + path + "' for reading");
  in.exceptions(std::ios::badbit | std::ios::failbit);
  try {
    uint32_t flag = 0;
    rd(in, flag);
    h->regular = false;
    if (!flag) {
      read_regular_body(h, in);
      return;
    }
    h->method = 3;
    rd(in, h->n);
    rd(in, h->elem);
    rd(in, h->off_l0);
    rd(in, h->off_data);
    rd(in, h->maxlevel);
    rd(in, h->enter);
    rd(in, h->maxM);
    rd(in, h->maxM0);
    rd(in, h->func);
    rd(in, h->method);
    if (h->func < kL2Sqr16 || h->func > kLInf)
      fail(NMSAMD_ERUNTIME, "Unknown distance function code: " + std::to_string(h->func));
    if (h->off_data != 0 || h->off_l0 < 16 || (h->off_l0 - 16) % 4 != 0 ||
        h->elem != h->off_l0 + 4 * (h->maxM0 + 1))
      fail(NMSAMD_ERUNTIME, "inconsistent optimized-index header");
    if (h->maxM0 > (uint64_t)kMaxM0 || h->maxM > (uint64_t)kMaxM0)
      fail(NMSAMD_ERUNTIME, "maxM0 " + std::to_string(h->maxM0) + " exceeds the device limit of " +
                                std::to_string(kMaxM0));
    h->dim = (int)((h->off_l0 - 16) / 4);
    h->block.resize((size_t)h->n * h->elem);
    if (!h->block.empty()) in.read(h->block.data(), h->block.size());
    h->sizemass.assign(h->n, 0);
    h->upper.clear();
    h->upper_off.assign(h->n, -1);
    for (uint32_t i = 0; i < h->n; ++i) {
      uint32_t sm = 0;
      rd(in, sm);
      h->sizemass[i] = sm;
      if (sm) {
        if (sm % (4 * (h->maxM + 1)) != 0) fail(NMSAMD_ERUNTIME, "corrupt upper-level list size");
        h->upper_off[i] = (int32_t)h->upper.size();
        size_t base = h->upper.size();
        h->upper.resize(base + sm / 4);
        in.read(reinterpret_cast<char*>(h->upper.data() + base), sm);
      }
    }
  } catch (const std::ios_base::failure& e) {
    fail(NMSAMD_EIO, std::string("error reading index '") + path + "': " + e.what());
  }
}

// Validates the host copy (ids in range, list counts within capacity, the
// entry point on the top level, the space's distance function) and notes
// whether any level-0 list repeats an id.
void check_index(nmsamd_index* h) {
  if (h->maxM0 > (uint64_t)kMaxM0 || h->maxM > (uint64_t)kMaxM0)
    fail(NMSAMD_ERUNTIME, "maxM0 " + std::to_string(h->maxM0) + " exceeds the device limit of " +
                              std::to_string(kMaxM0));
  if (h->n > 0 && h->enter >= h->n) fail(NMSAMD_ERUNTIME, "entry point out of range");
  if (h->expected_func == 1 && h->func != kL2Sqr16 && h->func != kL2Sqr && h->func != kSpaceL2)
    fail(NMSAMD_ERUNTIME, "index distance function does not match space l2");
  if (h->expected_func == kNormCos && h->func != kNormCos && h->func != kSpaceCos)
    fail(NMSAMD_ERUNTIME, "index distance function does not match space cosinesimil");
  if (h->expected_func == kNegDot && h->func != kNegDot && h->func != kSpaceNegDot)
    fail(NMSAMD_ERUNTIME, "index distance function does not match space negdotprod");
  if (h->expected_func == kL1 && h->func != kL1)
    fail(NMSAMD_ERUNTIME, "index distance function does not match space l1");
  if (h->expected_func == kLInf && h->func != kLInf)
    fail(NMSAMD_ERUNTIME, "index distance function does not match space linf");
  const int64_t n = h->n;
  bool dups = false;
  std::vector<int32_t> tmp;
  for (int64_t i = 0; i < n; ++i) {
    const char* e = h->block.data() + i * h->elem;
    int32_t cnt;
    std::memcpy(&cnt, e + h->off_l0, 4);
    if (cnt < 0 || (uint64_t)cnt > h->maxM0) fail(NMSAMD_ERUNTIME, "corrupt level-0 list count");
    const int32_t* ids = reinterpret_cast<const int32_t*>(e + h->off_l0 + 4);
    tmp.assign(ids, ids + cnt);
    for (int32_t v : tmp)
      if (v < 0 || v >= n) fail(NMSAMD_ERUNTIME, "level-0 neighbour id out of range");
    std::sort(tmp.begin(), tmp.end());
    if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end()) dups = true;
  }
  // a node's level is the number of its upper-level blocks
  auto level_of = [&](int64_t i) {
    return h->upper_off[i] < 0 ? 0 : (int)(h->sizemass[i] / (4 * (h->maxM + 1)));
  };
  for (size_t i = 0; i < h->upper_off.size(); ++i) {
    if (h->upper_off[i] < 0) continue;
    const int levels = level_of((int64_t)i);
    for (int L = 0; L < levels; ++L) {
      const int32_t* lst = h->upper.data() + h->upper_off[i] + (h->maxM + 1) * L;
      if (lst[0] < 0 || (uint64_t)lst[0] > h->maxM)
        fail(NMSAMD_ERUNTIME, "corrupt upper-level list count");
      for (int j = 1; j <= lst[0]; ++j) {
        if (lst[j] < 0 || lst[j] >= n) fail(NMSAMD_ERUNTIME, "upper-level neighbour id out of range");
        // the greedy descent reads a level-(L+1) neighbour's own list at that
        // level: HNSW links a node only to nodes present at the link's level
        if (level_of(lst[j]) < L + 1)
          fail(NMSAMD_ERUNTIME, "upper-level neighbour " + std::to_string(lst[j]) + " of node " +
                                    std::to_string(i) + " has no list at level " +
                                    std::to_string(L + 1));
      }
    }
  }
  if (n > 0 && h->maxlevel > 0) {
    const int lev_enter = h->upper_off[h->enter] < 0
                              ? 0
                              : (int)(h->sizemass[h->enter] / (4 * (h->maxM + 1)));
    if (lev_enter < h->maxlevel) fail(NMSAMD_ERUNTIME, "entry point level below maxlevel");
  }
  h->has_dups = dups;
}

// Lays the checked host copy out for the device (hnsw_device.h) and uploads it.
void upload_index(nmsamd_index* h) {
  const int64_t n = h->n;
  h->dstride = (h->dim + 3) / 4 * 4;
  h->l0stride = ((int)h->maxM0 + 31) / 32 * 32;
  std::vector<float> vec((size_t)std::max<int64_t>(n, 1) * h->dstride, 0.f);
  std::vector<int32_t> l0((size_t)std::max<int64_t>(n, 1) * h->l0stride, -1);
  std::vector<int32_t> ext((size_t)std::max<int64_t>(n, 1), 0);
  for (int64_t i = 0; i < n; ++i) {
    const char* e = h->block.data() + i * h->elem;
    std::memcpy(&ext[i], e, 4);
    std::memcpy(&vec[i * h->dstride], e + 16, 4 * (size_t)h->dim);
    int32_t cnt;
    std::memcpy(&cnt, e + h->off_l0, 4);
    std::memcpy(&l0[i * h->l0stride], e + h->off_l0 + 4, 4 * (size_t)cnt);
  }
  ensure_device(h);
  hipStream_t s = h->stream;
  h->d_vec.upload(vec.data(), vec.size(), s);
  h->d_l0.upload(l0.data(), l0.size(), s);
  std::vector<int32_t> upoff(h->upper_off.begin(), h->upper_off.end());
  if (upoff.empty()) upoff.push_back(-1);
  h->d_upoff.upload(upoff.data(), upoff.size(), s);
  std::vector<int32_t> up = h->upper;
  if (up.empty()) up.push_back(0);
  h->d_up.upload(up.data(), up.size(), s);
  h->d_ext.upload(ext.data(), ext.size(), s);
  hip_check(hipStreamSynchronize(s), "upload");
  h->d_vec16.release();
  h->d_vec8.release();
  h->d_scale8.release();
  h->d_xnorm.release();
  h->vec_bounds_known = false;
  h->d_visited.release();
  h->d_epoch.release();
  h->slots = 0;
  h->on_device = true;
}

// Entry points that search: an index must be loaded or built; a built one
// reaches the device on its first use.
void require_index(nmsamd_index* h) {
  if (!h->loaded) fail(NMSAMD_EINVAL, "Must call createIndex or loadIndex before this method");
  if (!h->on_device) upload_index(h);
}

// Visited-mark arrays for `slots` resident waves.
void ensure_slots(nmsamd_index* h, int slots) {
  if (slots <= h->slots) return;
  h->vstride = ((int64_t)h->n + 1 + 255) / 256 * 256;
  h->d_visited.release();
  h->d_epoch.release();
  h->d_visited.ensure((size_t)slots * h->vstride);
  h->d_epoch.ensure((size_t)slots);
  hip_check(hipMemsetAsync(h->d_visited.p, 0, (size_t)slots * h->vstride, h->stream), "memset");
  hip_check(hipMemsetAsync(h->d_epoch.p, 0, (size_t)slots, h->stream), "memset");
  h->slots = slots;
  h->slots_bits = false;
}

// Visited marks of a beam search as bits (a query clears its slot's n / 8
// bytes when it starts) where that clearing is small next to the query's own
// mark probes (about ef * maxM0), else epoch bytes (no clearing).  The exact
// searches (V1Merge, SearchOld) always take epoch bytes.
bool use_bit_marks(const nmsamd_index* h, int max_ef, bool beam) {
  if (!beam) return false;
  if (h->visited_marks) return h->visited_marks == 2;
  return ((int64_t)h->n + 1) / 8 <= (int64_t)8 * std::max(max_ef, 1) * std::max<int64_t>(h->maxM0, 1);
}

// The reduced-precision copy the current storage traverses, converted on the
// first use.  The conversion finishes before this returns: searches and
// serving lanes on other streams read the copy right after.
void ensure_lowp(nmsamd_index* h, bool need16 = false) {
  bool converted = false;
  if ((h->storage == kBF16 || need16) && !h->d_vec16.p) {
    h->d_vec16.ensure((size_t)std::max<int64_t>(h->n, 1) * h->dstride);
    hip_check(launch_convert_lowp(h->d_vec.p, h->d_vec16.p, nullptr, nullptr, h->n, h->dim,
                                  h->dstride, kBF16, h->stream), "convert bf16");
    converted = true;
  }
  if (h->storage == kFP8 && !h->d_vec8.p) {
    h->d_vec8.ensure((size_t)std::max<int64_t>(h->n, 1) * h->dstride);
    h->d_scale8.ensure((size_t)std::max<int64_t>(h->n, 1));
    hip_check(launch_convert_lowp(h->d_vec.p, nullptr, h->d_vec8.p, h->d_scale8.p, h->n, h->dim,
                                  h->dstride, kFP8, h->stream), "convert fp8");
    converted = true;
  }
  if (converted) hip_check(hipStreamSynchronize(h->stream), "reduced-precision conversion");
}

DevIndex dev_index(nmsamd_index* h) {
  DevIndex ix{};
  ix.vec = h->d_vec.p;
  ix.vec16 = h->d_vec16.p;
  ix.vec8 = h->d_vec8.p;
  ix.scale8 = h->d_scale8.p;
  ix.l0 = h->d_l0.p;
  ix.up_off = h->d_upoff.p;
  ix.up = h->d_up.p;
  ix.ext_id = h->d_ext.p;
  ix.visited = h->d_visited.p;
  ix.epoch = h->d_epoch.p;
  ix.n = h->n;
  ix.vstride = h->vstride;
  ix.dim = h->dim;
  ix.dstride = h->dstride;
  ix.l0stride = h->l0stride;
  ix.maxM = (int)h->maxM;
  ix.maxM0 = (int)h->maxM0;
  ix.maxlevel = h->maxlevel;
  ix.enter = (int)h->enter;
  ix.func = h->func;
  ix.has_dups = h->has_dups ? 1 : 0;
  return ix;
}

// Beam mode on the cooperative kernel (one query per block of kCoopWaves
// waves): forced by setBeamWaves(8), and the only kernel for beam widths
// above 16.
bool use_coop(const nmsamd_index* h) { return h->beam > 0 && (h->beam_waves == 8 || h->beam > 16); }
// Two cooperative blocks per CU (kModeBeamCoop2) where LDS holds two queries,
// unless setBeamGather(2) asks for the wide distance path of one per CU.
bool coop_two_per_cu(const nmsamd_index* h, size_t fit) { return fit >= 2 && h->beam_gather != 2; }
// The BF16 screen where it applies: the cooperative kernel on FP32 vectors,
// squared l2 of an optimized index (SearchArgs::screen).
int screen_mode(const nmsamd_index* h) {
  return (h->storage == kF32 && (h->func == kL2Sqr16 || h->func == kL2Sqr)) ? h->screen : 0;
}

// Neighbour entries one search step holds in LDS: beam mode takes up to
// kMaxStepEntries / maxM0 expansions per step (kMaxCoopEntries on the
// cooperative kernel), W * maxM0 entries rounded up to whole rows of 64; the
// other modes expand one list of <= 64 entries.
constexpr int kMaxStepEntries = 512;
constexpr int kMaxCoopEntries = 2048;
int step_entries(const nmsamd_index* h) {
  if (h->beam <= 0) return 128;
  const int maxe = use_coop(h) ? kMaxCoopEntries : kMaxStepEntries;
  const int w = std::max(1, std::min(h->beam, maxe / (int)h->maxM0));
  return std::max(128, (w * (int)h->maxM0 + 63) / 64 * 64);
}

// sorted array + query (FP32, then BF16 and its norm: 8 * dstride bytes) + two
// scratch regions of 8 * ent bytes; beam mode
// adds the merge's 16-byte record per 64 positions, the claimed next picks
// (416 bytes) and the two-wave command words (96 bytes)
int lds_wave_bytes(int cap, int dstride, int ent, bool beam) {
  int b = ((cap + 1) & ~1) * 8 + dstride * 8 + 16 * ent;
  if (beam) b += (cap / 64 + 2) * 16 + 416 + 96;
  return (b + 15) / 16 * 16;
}

// LDS a two-wave block reserves: more than a quarter of the CU's, so that a
// fourth block never fits, plus the second items buffer of the deferred
// rebuild (8 * ent bytes) where that keeps the CU's number of queries.
int pair_lds(int lds, int ent, bool* defer) {
  const int floor = (int)(kLdsPerCu / 4) + 16;
  const int plain = std::max(lds, floor);
  const int with2 = std::max(lds + 8 * ent, floor);
  *defer = kLdsPerCu / (size_t)with2 == kLdsPerCu / (size_t)plain;
  return *defer ? with2 : plain;
}

// The handle's launches share its work counter, error word, visited slots,
// epochs and events: a launch on another stream than the last one starts only
// after the last has finished (its end is ev1; the device entry point returns
// without waiting).  The caller records ev1 after its own launch.
void follow_last_launch(nmsamd_index* h, hipStream_t s) {
  if (h->last_stream && h->last_stream != s)
    hip_check(hipStreamWaitEvent(s, h->ev1, 0), "hipStreamWaitEvent");
  h->last_stream = s;
}

// Device memory a launch may use instead of the handle's own buffers: the
// work counter and error word, the slots' epochs and visited marks and the
// SearchOld heaps are laid out in [base, base + bytes).
struct ExtScratch {
  unsigned char* base;
  size_t bytes;
};

// Scratch of nmsamd_knn_query_batch_device_noalloc (256 MiB per GPU).
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// The address of GPU `device`'s copy of `scratch`, taken once per device with
// that device current.
unsigned char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) fail(NMSAMD_EDEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device) hip_check(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!p) fail(NMSAMD_EDEVICE, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

// Launches the search of nq queries whose inputs/outputs are device pointers;
// k_of / ef_of are optional device arrays whose maxima are max_k / max_ef.
// With `ext`, the launch allocates nothing: its scratch comes from ext (the
// resident slots are as many as fit) and no counters are collected.
void run_search(nmsamd_index* h, const float* d_q, int64_t nq, int k, int ef, const int32_t* k_of,
                const int32_t* ef_of, int max_k, int max_ef, int32_t* ids, float* dists,
                int32_t* counts, hipStream_t s, const ExtScratch* ext = nullptr) {
  if (!h->loaded) fail(NMSAMD_EINVAL, "Must call createIndex or loadIndex before this method");
  // Hnsw::Search (hnsw.cc:724): SearchOld for algoType=old or hybrid with ef
  // >= 1000; beam mode replaces V1Merge and the hybrid switch for the batch
  const bool beam = h->beam > 0 && h->algo != kOld;
  const bool any_old = !beam && (h->algo == kOld || (h->algo == kHybrid && max_ef >= 1000));
  int cap = std::max(1, std::max(max_ef, max_k));
  if (any_old) cap = std::max(cap, max_ef + 1 + max_k);
  if (h->regular && h->storage != kF32)
    fail(NMSAMD_EINVAL, "reduced-precision traversal needs an optimized index (a regular one is "
                        "searched with the space's own FP32 distance)");
  const bool rerank = h->storage != kF32 && h->rerank_mult > 0;
  if (rerank) {  // staging room for the re-sorted candidates
    const int r = std::min(std::max(max_ef, max_k), std::max(max_k, h->rerank_mult * max_k));
    cap = std::max(cap, std::max(max_ef, max_k) + r);
  }
  const int ent = beam ? step_entries(h) : 128;
  const bool coop = beam && use_coop(h);
  // the cooperative kernel takes bit marks (its visited probe is an atomic OR)
  const bool vbits = coop || use_bit_marks(h, max_ef, beam);
  const int lds = lds_wave_bytes(cap, h->dstride, ent, beam);
  if ((size_t)lds > kLdsPerCu)
    fail(NMSAMD_EINVAL, "efSearch / k too large: the search array of " + std::to_string(cap) +
                            " items exceeds the LDS of one wave");
  const int screen = coop ? screen_mode(h) : 0;
  if (ext) {
    if (((h->storage == kBF16 || screen) && !h->d_vec16.p) || (h->storage == kFP8 && !h->d_vec8.p))
      fail(NMSAMD_EINVAL, "the reduced-precision copy of the vectors is made by the first "
                          "search in that precision: run one through another entry point first");
  } else {
    ensure_lowp(h, screen != 0);
  }
  // queries per block and resident blocks per CU from the LDS footprint; the
  // grid is persistent: as many resident queries as the device holds, at most
  // one per query.  Beam batches take the beam-only kernels: where the LDS
  // footprint leaves room for fewer than two queries per SIMD, each query gets
  // two waves (kModeBeamPair, one query per block, at most two waves per
  // SIMD), else one.
  // With one wave per query, the wide beam kernel (registers for one wave
  // per SIMD) is taken where the footprint limits residency to that.
  // Two waves per query run at most three queries per CU, the shape they are
  // measured and tested in (variants at four queries per CU faulted the GPU
  // on a 10,000-point index with the cause not found), and are chosen
  // automatically only where LDS already limits a CU to three queries.  The
  // limit is the hardware's, not only the grid's: a two-wave block reserves
  // more than a quarter of the CU's LDS, so a fourth block never fits.
  // A cooperative block takes the CU's register file at two waves per SIMD
  // (one query per CU), or half of it where LDS holds two queries
  // (kModeBeamCoop2).
  const size_t fit = kLdsPerCu / lds;
  const bool coop2 = coop && coop_two_per_cu(h, fit);
  const bool pair = beam && !coop && (h->beam_waves == 2 || (h->beam_waves == 0 && fit <= 3));
  bool defer = false;
  const int lds_launch = pair ? pair_lds(lds, ent, &defer) : lds;
  const int wpb = (pair || coop) ? 1 : (int)std::max<size_t>(1, std::min<size_t>(kWavesPerBlock, fit));
  const int per_cu = coop ? (coop2 ? 2 : 1) : (int)std::max<size_t>(
      1, std::min<size_t>(pair ? 3 : 8, kLdsPerCu / ((size_t)lds_launch * wpb)));
  // the one-wave kernel: wide (every vector load of a round in flight, the
  // register file of one wave per SIMD) where the footprint already holds a
  // CU to four waves, or as forced by setBeamGather; forced wide, the grid
  // keeps to the four waves per CU it can hold
  const bool wide = !pair && !coop &&
                    (h->beam_gather == 2 || (h->beam_gather == 0 && wpb * per_cu <= 4));
  const int per_cu_run = wide ? std::max(1, std::min(per_cu, 4 / wpb)) : per_cu;
  const int mode = !beam ? kModeAll
                         : (coop ? (coop2 ? kModeBeamCoop2 : kModeBeamCoop)
                                 : (pair ? kModeBeamPair : (wide ? kModeBeamWide : kModeBeam)));
  int64_t max_blocks = (int64_t)h->cu_count * per_cu_run;
  const int64_t heap_cap = any_old ? std::max<int64_t>(65536, 64 * (int64_t)max_ef) : 0;
  // a slot of the external scratch holds n + 1 epoch bytes, or with bit marks
  // the (n + 128) / 128 16-byte words a query clears (hnsw_search_impl.h)
  const int64_t vstride = vbits ? (((int64_t)h->n + 128) / 128 * 16 + 255) / 256 * 256
                                : ((int64_t)h->n + 1 + 255) / 256 * 256;
  // visited marks take n bytes per slot: keep them under ~16 GiB, or within
  // the external scratch (256 B of counters, the epochs, then per slot its
  // marks and SearchOld heap)
  int64_t slot_cap = std::max<int64_t>(wpb, ((int64_t)16 << 30) / std::max<int64_t>(h->n, 1));
  if (ext) {
    const int64_t per_slot = vstride + heap_cap * 8 + 1;
    slot_cap = ((int64_t)ext->bytes - 512) / per_slot;
    if (slot_cap < wpb)
      fail(NMSAMD_EINVAL, "the no-allocation scratch of " + std::to_string(ext->bytes >> 20) +
                              " MiB holds no search slot for an index of " + std::to_string(h->n) +
                              " points" + (any_old ? " with the SearchOld heap" : ""));
  }
  max_blocks = std::max<int64_t>(1, std::min<int64_t>(max_blocks, slot_cap / wpb));
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((nq + wpb - 1) / wpb, max_blocks));
  const int slots = blocks * wpb;
  int* work = nullptr;
  DevIndex ix = dev_index(h);
  float* old_key = nullptr;
  int32_t* old_id = nullptr;
  follow_last_launch(h, s);
  if (ext) {
    const int64_t ep_bytes = ((int64_t)slots + 255) / 256 * 256;
    work = reinterpret_cast<int*>(ext->base);
    ix.epoch = ext->base + 256;
    ix.visited = ext->base + 256 + ep_bytes;
    ix.vstride = vstride;
    old_key = reinterpret_cast<float*>(ix.visited + (int64_t)slots * vstride);
    old_id = reinterpret_cast<int32_t*>(old_key + (int64_t)slots * heap_cap);
    // the scratch keeps the last call's contents (possibly another index's
    // layout): counters, epochs and marks start from zero
    const int64_t clear = 256 + ep_bytes + (int64_t)slots * vstride;
    hip_check(launch_fill_i32(work, 0, (int)(clear / 4), s), "fill");
  } else {
    ensure_slots(h, slots);
    ix = dev_index(h);
    // bit marks left in the slots would read as stale epochs: clear them
    // before an epoch-byte launch (an epoch never equals a zero byte)
    if (!vbits && h->slots_bits) {
      const size_t words = (size_t)h->slots * h->vstride / 4;
      for (size_t o = 0; o < words; o += (size_t)1 << 30)
        hip_check(launch_fill_i32(reinterpret_cast<int*>(h->d_visited.p) + o, 0,
                                  (int)std::min(words - o, (size_t)1 << 30), s), "fill");
      h->slots_bits = false;
    }
    if (vbits) h->slots_bits = true;
    if (any_old) {
      h->d_oldk.ensure((size_t)slots * heap_cap);
      h->d_oldi.ensure((size_t)slots * heap_cap);
    }
    old_key = h->d_oldk.p;
    old_id = h->d_oldi.p;
    h->d_work.ensure(2);
    work = h->d_work.p;
    hip_check(launch_fill_i32(work, 0, 2, s), "fill");
  }
  const bool collect = h->collect && !ext;
  if (collect) {
    h->d_stats.ensure(1);
    hip_check(launch_fill_i32(reinterpret_cast<int*>(h->d_stats.p), 0,
                              sizeof(SearchStats) / sizeof(int), s), "fill");
  }
  SearchArgs a{};
  a.ix = ix;
  a.queries = d_q;
  a.nq = nq;
  a.k_of = k_of;
  a.ef_of = ef_of;
  a.order = h->order_dev;
  a.k = k;
  a.ef = ef;
  a.cap = cap;
  a.out_stride = max_k;
  a.out_ids = ids;
  a.out_dists = dists;
  a.out_counts = counts;
  a.out_internal = nullptr;
  a.work = work;
  a.err = work + 1;
  a.stats = collect ? h->d_stats.p : nullptr;
  if (collect) h->d_qticks.ensure((size_t)nq);
  a.q_ticks = collect ? h->d_qticks.p : nullptr;
  h->last_nq = nq;
  a.storage = h->storage;
  a.rerank = rerank ? h->rerank_mult : 0;
  a.beam = h->beam;
  a.ent = ent;
  a.algo = h->algo == kOld ? kAlgoOld : (h->algo == kV1Merge ? kAlgoV1Merge : kAlgoHybrid);
  a.old_heap_key = old_key;
  a.old_heap_id = old_id;
  a.old_heap_cap = heap_cap;
  a.defer = defer ? 1 : 0;
  a.vbits = vbits ? 1 : 0;
  a.screen = screen;
  hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
  hip_check(launch_search(a, blocks, wpb, lds_launch, mode, s), "search kernel launch");
  hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
  h->last_launch[0] = mode;
  h->last_launch[1] = blocks;
  h->last_launch[2] = wpb;
  h->last_launch[3] = lds_launch;
  h->last_launch[4] = ent;
  h->last_err = work + 1;
  h->last_collect = collect;
  h->pending = true;
}

void finish_stats(nmsamd_index* h) {
  if (!h->pending) return;
  h->pending = false;
  int err = 0;
  hip_check(hipMemcpy(&err, h->last_err, sizeof(int), hipMemcpyDeviceToHost), "status");
  if (err & 2)
    fail(NMSAMD_ERUNTIME, "internal error: corrupted search state (a node id outside the index, or an item expanded twice)");
  if (err & 4)
    fail(NMSAMD_ERUNTIME, "internal error: the BF16 screen's bound failed against an FP32 distance");
  if (err)
    fail(NMSAMD_ERUNTIME, "SearchOld candidate heap overflow (more than 64*ef admitted items)");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
  h->last_ms = ms;
  if (h->last_collect)
    hip_check(hipMemcpy(&h->last_stats, h->d_stats.p, sizeof(SearchStats), hipMemcpyDeviceToHost),
              "stats");
}


// AnyParams from "key=value" strings (params.h:45-120): one '=' per string,
// no repeated key.
std::vector<std::pair<std::string, std::string>> parse_params(const char* const* kv, int n) {
  std::vector<std::pair<std::string, std::string>> params;
  for (int i = 0; i < n; ++i) {
    const std::string s = kv[i] ? kv[i] : "";
    const size_t eq = s.find('=');
    if (eq == std::string::npos || s.find('=', eq + 1) != std::string::npos)
      fail(NMSAMD_ERUNTIME, "Wrong format of an argument: '" + s +
                                "' should be in the format: <Name>=<Value>");
    const std::string name = s.substr(0, eq);
    for (auto& p : params)
      if (p.first == name) fail(NMSAMD_ERUNTIME, "Duplicate parameter: " + name);
    params.emplace_back(name, s.substr(eq + 1));
  }
  return params;
}

// AnyParamManager::ConvertStrToValue (params.h:295-305): the whole string
// must convert; the message names the type as typeid does (m, i, d).
size_t to_size(const std::string& v) {
  std::stringstream ss(v);
  size_t out = 0;
  if (!(ss >> out) || !ss.eof() || v.find('-') != std::string::npos)
    fail(NMSAMD_ERUNTIME, "Failed to convert value '" + v + "' from type: m");
  return out;
}
int to_int(const std::string& v) {
  std::stringstream ss(v);
  int out = 0;
  if (!(ss >> out) || !ss.eof()) fail(NMSAMD_ERUNTIME, "Failed to convert value '" + v + "' from type: i");
  return out;
}
double to_double(const std::string& v) {
  std::stringstream ss(v);
  double out = 0;
  if (!(ss >> out) || !ss.eof()) fail(NMSAMD_ERUNTIME, "Failed to convert value '" + v + "' from type: d");
  return out;
}

// Refuses to replace or free an index that open servers still search.
void require_no_servers(const nmsamd_index* h) {
  if (h->servers > 0)
    fail(NMSAMD_EINVAL, "close the servers opened on this index first (" +
                            std::to_string(h->servers) + " open)");
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return NMSAMD_OK;
  } catch (const Status& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return NMSAMD_EINVAL;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NMSAMD_ERUNTIME;
  }
}

}  // namespace

extern "C" {

const char* nmsamd_last_error(void) { return g_err.c_str(); }

int nmsamd_init(const char* space, const char* method, nmsamd_index** out) {
  return guarded([&] {
    if (!out) fail(NMSAMD_EINVAL, "null output handle");
    const std::string sp = space ? space : "cosinesimil";
    const std::string me = method ? method : "hnsw";
    auto h = std::make_unique<nmsamd_index>();
    if (sp == "l2") h->expected_func = 1;
    else if (sp == "cosinesimil") h->expected_func = kNormCos;
    else if (sp == "negdotprod") h->expected_func = kNegDot;
    else if (sp == "l1") h->expected_func = kL1;
    else if (sp == "linf") h->expected_func = kLInf;
    else fail(NMSAMD_ERUNTIME, "It looks like the space " + sp +
                                   " is not defined for the distance type : FLOAT");
    if (me != "hnsw")
      fail(NMSAMD_ERUNTIME, "It looks like the method " + me +
                                " is not defined for the distance type : FLOAT (only hnsw is available)");
    h->space = sp;
    *out = h.release();
  });
}

int nmsamd_free(nmsamd_index* h) {
  return guarded([&] {
    if (h) require_no_servers(h);
    delete h;
  });
}

int nmsamd_load_index(nmsamd_index* h, const char* filename, int load_data) {
  return guarded([&] {
    if (!h || !filename) fail(NMSAMD_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    require_no_servers(h);
    h->loaded = false;
    if (load_data) {  // without load_data the points added so far stay (nmslib.cc:103-107)
      h->data_objs.clear();
      const std::string dat = std::string(filename) + ".dat";
      std::ifstream in(dat, std::ios::binary);
      if (!in) fail(NMSAMD_EIO, "Cannot open file '" + dat + "' for reading");
      in.exceptions(std::ios::badbit | std::ios::failbit);
      try {
        uint64_t qty = 0;
        rd(in, qty);
        for (uint64_t i = 0; i < qty; ++i) {
          uint64_t sz = 0;
          rd(in, sz);
          std::vector<char> obj(sz);
          in.read(obj.data(), sz);
          h->data_objs.push_back(std::move(obj));
        }
      } catch (const std::ios_base::failure& e) {
        fail(NMSAMD_EIO, "error reading '" + dat + "': " + e.what());
      }
    }
    h->on_device = false;
    read_index_file(h, filename);
    check_index(h);
    int devices = 0;  // upload now when a device is visible, else on the first search
    if (hipGetDeviceCount(&devices) == hipSuccess && devices > 0) upload_index(h);
    h->loaded = true;
    h->ef = 20;  // ResetQueryTimeParams (nmslib.cc:110-112)
    h->algo = kHybrid;
  });
}

int nmsamd_save_index(nmsamd_index* h, const char* filename, int save_data) {
  return guarded([&] {
    if (!h || !filename) fail(NMSAMD_EINVAL, "null argument");
    if (!h->loaded) fail(NMSAMD_EINVAL, "Must call createIndex or loadIndex before this method");
    if (save_data) {
      const std::string dat = std::string(filename) + ".dat";
      std::ofstream out(dat, std::ios::binary);
      if (!out) fail(NMSAMD_EIO, "Cannot open file '" + dat + "' for writing");
      wr(out, (uint64_t)h->data_objs.size());
      for (auto& o : h->data_objs) {
        wr(out, (uint64_t)o.size());
        out.write(o.data(), o.size());
      }
    }
    std::ofstream out(filename, std::ios::binary);
    if (!out) fail(NMSAMD_EIO, std::string("Cannot open file '") + filename + "' for writing");
    if (h->regular) {
      write_regular(h, out);
      if (!out) fail(NMSAMD_EIO, std::string("error writing '") + filename + "'");
      return;
    }
    wr(out, (uint32_t)1);
    wr(out, h->n);
    wr(out, h->elem);
    wr(out, h->off_l0);
    wr(out, h->off_data);
    wr(out, h->maxlevel);
    wr(out, h->enter);
    wr(out, h->maxM);
    wr(out, h->maxM0);
    wr(out, h->func);
    wr(out, h->method);
    out.write(h->block.data(), h->block.size());
    for (uint32_t i = 0; i < h->n; ++i) {
      wr(out, h->sizemass[i]);
      if (h->sizemass[i])
        out.write(reinterpret_cast<const char*>(h->upper.data() + h->upper_off[i]), h->sizemass[i]);
    }
    if (!out) fail(NMSAMD_EIO, std::string("error writing '") + filename + "'");
  });
}

int nmsamd_add_data_point_batch(nmsamd_index* h, const float* data, int64_t n, int64_t dim,
                                const int32_t* ids, int64_t* out_count) {
  return guarded([&] {
    if (!h || (n > 0 && !data)) fail(NMSAMD_EINVAL, "null argument");
    if (n < 0) fail(NMSAMD_EINVAL, "the number of points must be non-negative");
    if (dim <= 0) fail(NMSAMD_EINVAL, "data points need a positive dimension");
    std::lock_guard<std::mutex> g(h->mu);
    // Object buffer (object.h:40-104): i32 id, i32 label (-1), u64 datalength, data
    const uint64_t len = 4 * (uint64_t)dim;
    for (int64_t i = 0; i < n; ++i) {
      std::vector<char> obj(16 + len);
      const int32_t id = ids ? ids[i] : (int32_t)i;
      const int32_t label = -1;
      std::memcpy(obj.data(), &id, 4);
      std::memcpy(obj.data() + 4, &label, 4);
      std::memcpy(obj.data() + 8, &len, 8);
      std::memcpy(obj.data() + 16, data + i * dim, len);
      h->data_objs.push_back(std::move(obj));
    }
    if (out_count) *out_count = n;
  });
}

// Data point `pos` as a dense float vector (IndexWrapper::at, nmslib.cc:348).
const std::vector<char>& data_point(const nmsamd_index* h, int64_t pos) {
  if (pos < 0 || (uint64_t)pos >= h->data_objs.size())
    fail(NMSAMD_EINVAL, "data point position " + std::to_string(pos) + " out of range");
  const auto& o = h->data_objs[pos];
  uint64_t dl = 0;
  if (o.size() >= 16) std::memcpy(&dl, o.data() + 8, 8);
  if (o.size() < 16 || dl + 16 != o.size() || dl % 4 != 0)
    fail(NMSAMD_ERUNTIME, "data point " + std::to_string(pos) + " is not a dense float vector");
  return o;
}

int nmsamd_get_data_point(nmsamd_index* h, int64_t pos, float* out, int64_t cap, int64_t* dim) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    const auto& o = data_point(h, pos);
    const int64_t d = (int64_t)(o.size() - 16) / 4;
    if (dim) *dim = d;
    if (out) std::memcpy(out, o.data() + 16, 4 * (size_t)std::min(d, std::max<int64_t>(cap, 0)));
  });
}

int nmsamd_get_distance(nmsamd_index* h, int64_t pos1, int64_t pos2, float* out) {
  return guarded([&] {
    if (!h || !out) fail(NMSAMD_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(h->mu);
    const auto& a = data_point(h, pos1);
    const auto& b = data_point(h, pos2);
    if (a.size() != b.size()) fail(NMSAMD_ERUNTIME, "data points differ in dimension");
    const int space = h->expected_func == 1 ? (int)kBuildL2 : h->expected_func;
    *out = space_distance(space, reinterpret_cast<const float*>(a.data() + 16),
                          reinterpret_cast<const float*>(b.data() + 16), (a.size() - 16) / 4);
  });
}

int nmsamd_vector_distance(nmsamd_index* h, const float* a, const float* b, int64_t dim,
                           float* out) {
  return guarded([&] {
    if (!h || !a || !b || !out) fail(NMSAMD_EINVAL, "null argument");
    if (dim <= 0) fail(NMSAMD_EINVAL, "vectors need a positive dimension");
    const int space = h->expected_func == 1 ? (int)kBuildL2 : h->expected_func;
    *out = space_distance(space, a, b, (size_t)dim);
  });
}

int nmsamd_create_index(nmsamd_index* h, const char* const* kv, int nkv, int print_progress) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    require_no_servers(h);
    // Hnsw::CreateIndex parameters and defaults (hnsw.cc:187-205)
    BuildParams p;
    bool has_maxM = false, has_maxM0 = false, has_mult = false;
    int skip_optimized = 0;
    p.threads = std::max(1u, std::thread::hardware_concurrency());
    for (auto& kvp : parse_params(kv, nkv)) {
      const std::string& k = kvp.first;
      const std::string& v = kvp.second;
      if (k == "M") p.M = to_size(v);
      else if (k == "maxM") { p.maxM = to_size(v); has_maxM = true; }
      else if (k == "maxM0") { p.maxM0 = to_size(v); has_maxM0 = true; }
      else if (k == "efConstruction") p.efConstruction = to_size(v);
      else if (k == "indexThreadQty") {
        p.threads = to_size(v);
        if (p.threads == 0) p.threads = std::max(1u, std::thread::hardware_concurrency());
      }
      else if (k == "mult") { p.mult = to_double(v); has_mult = true; }
      else if (k == "delaunay_type") p.delaunay = to_int(v);
      else if (k == "post") p.post = to_int(v);
      else if (k == "skip_optimized_index") skip_optimized = to_int(v);
      else if (k == "searchMethod") (void)to_int(v);
      else fail(NMSAMD_ERUNTIME, "Unknown parameters found!");
    }
    if (!has_maxM) p.maxM = p.M;
    if (!has_maxM0) p.maxM0 = p.M * 2;
    if (!has_mult) p.mult = 1 / std::log(1.0 * p.M);
    p.optimized = skip_optimized == 0;
    p.progress = print_progress != 0;
    if (p.maxM0 > (size_t)kMaxM0 || p.maxM > (size_t)kMaxM0)
      fail(NMSAMD_EINVAL, "maxM0 " + std::to_string(p.maxM0) + " exceeds the device limit of " +
                              std::to_string(kMaxM0));
    if (h->data_objs.empty()) fail(NMSAMD_EINVAL, "no data points: call addDataPointBatch first");
    const int space = h->expected_func == 1 ? (int)kBuildL2 : h->expected_func;
    BuiltIndex b;
    try {
      build_hnsw(space, h->data_objs, p, b);
    } catch (const std::invalid_argument& e) {
      fail(NMSAMD_EINVAL, e.what());
    }
    h->loaded = false;
    h->on_device = false;
    h->regular = !p.optimized;
    h->reg_M = p.M;
    h->file_maxlevel = b.maxlevel;
    h->n = b.n;
    h->elem = b.elem;
    h->off_l0 = b.off_l0;
    h->off_data = 0;
    h->maxlevel = b.maxlevel;
    h->enter = b.enter;
    h->maxM = b.maxM;
    h->maxM0 = b.maxM0;
    h->func = b.func;
    h->method = p.optimized ? 3 : 0;  // searchMethod: 3 for every optimized index (hnsw.cc:373)
    h->dim = (int)((b.off_l0 - 16) / 4);
    h->block = std::move(b.block);
    h->sizemass = std::move(b.sizemass);
    h->upper = std::move(b.upper);
    h->upper_off = std::move(b.upper_off);
    check_index(h);
    h->loaded = true;
    h->ef = 20;  // CreateIndex resets the query-time parameters (hnsw.cc:207)
    h->algo = kHybrid;
  });
}

int nmsamd_set_query_time_params(nmsamd_index* h, const char* const* kv, int n) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    const auto params = parse_params(kv, n);
    auto has = [&](const char* name) {
      for (auto& p : params) if (p.first == name) return true;
      return false;
    };
    if (has("ef") && has("efSearch"))
      fail(NMSAMD_ERUNTIME,
           "The user shouldn't specify parameters ef and efSearch at the same time (they are synonyms)");
    size_t ef = 20;
    Algo algo = kHybrid;
    for (auto& p : params) {
      if (p.first == "ef" || p.first == "efSearch") {
        ef = to_size(p.second);
      } else if (p.first == "searchMethod") {
        (void)to_int(p.second);
      } else if (p.first == "algoType") {
        const std::string t = to_lower(p.second);
        if (t == "v1merge") algo = kV1Merge;
        else if (t == "old") algo = kOld;
        else if (t == "hybrid") algo = kHybrid;
        else fail(NMSAMD_ERUNTIME, "algoType should be one of the following: old, v1merge");
      } else {
        fail(NMSAMD_ERUNTIME, "Unknown parameters found!");
      }
    }
    h->ef = ef;
    h->algo = algo;
  });
}

int nmsamd_knn_query_batch(nmsamd_index* h, const float* q, int64_t nq, int64_t dim, int32_t k,
                           int32_t num_threads, int32_t* ids, float* dists, int32_t* counts) {
  (void)num_threads;
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    require_index(h);
    if (nq < 0 || k < 0) fail(NMSAMD_EINVAL, "nq and k must be non-negative");
    if (dim != h->dim)
      fail(NMSAMD_EINVAL, "query dimension " + std::to_string(dim) + " != index dimension " +
                              std::to_string(h->dim));
    if (nq == 0) return;
    if (k == 0) {
      std::fill(counts, counts + nq, 0);
      return;
    }
    hipStream_t s = h->stream;
    h->d_q.upload(q, (size_t)nq * dim, s);
    h->d_ids.ensure((size_t)nq * k);
    h->d_dists.ensure((size_t)nq * k);
    h->d_counts.ensure((size_t)nq);
    run_search(h, h->d_q.p, nq, k, (int)h->ef, nullptr, nullptr, k, (int)h->ef, h->d_ids.p,
               h->d_dists.p, h->d_counts.p, s);
    hip_check(hipMemcpyAsync(ids, h->d_ids.p, sizeof(int32_t) * nq * k, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(dists, h->d_dists.p, sizeof(float) * nq * k, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(counts, h->d_counts.p, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "search");
    finish_stats(h);
  });
}

int nmsamd_knn_query_batch_device(nmsamd_index* h, const float* d_q, int64_t nq, int32_t k,
                                  int32_t* d_ids, float* d_dists, int32_t* d_counts, void* stream) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    require_index(h);
    if (nq < 0 || k <= 0) fail(NMSAMD_EINVAL, "nq must be non-negative and k positive");
    if (nq == 0) return;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    run_search(h, d_q, nq, k, (int)h->ef, nullptr, nullptr, k, (int)h->ef, d_ids, d_dists, d_counts, s);
  });
}

// The no-allocation variant of nmsamd_knn_query_batch_device: its scratch
// (work counter, error word, epochs, visited marks, SearchOld heaps) is this
// file-scope device array, one copy per GPU, held until the process ends.
// Host code addresses a GPU's copy only through the pointer
// hipGetSymbolAddress returns for it on that GPU (scratch_on()).
int nmsamd_knn_query_batch_device_noalloc(nmsamd_index* h, const float* d_q, int64_t nq,
                                          int32_t k, int32_t* d_ids, float* d_dists,
                                          int32_t* d_counts, void* stream) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->loaded) fail(NMSAMD_EINVAL, "Must call createIndex or loadIndex before this method");
    if (!h->on_device)
      fail(NMSAMD_EINVAL, "the index is not in device memory yet: the no-allocation entry point "
                          "does not upload it (any other search does)");
    if (nq < 0 || k <= 0) fail(NMSAMD_EINVAL, "nq must be non-negative and k positive");
    if (nq == 0) return;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    const ExtScratch ext{scratch_on(h->device), kNoAllocScratchBytes};
    run_search(h, d_q, nq, k, (int)h->ef, nullptr, nullptr, k, (int)h->ef, d_ids, d_dists,
               d_counts, s, &ext);
  });
}

int nmsamd_knn_query_batch_uneven(nmsamd_index* h, const float* q, int64_t nq, int64_t dim,
                                  const int32_t* k, const int32_t* ef, int32_t max_k,
                                  int32_t* ids, float* dists, int32_t* counts) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    require_index(h);
    if (dim != h->dim) fail(NMSAMD_EINVAL, "query dimension does not match the index");
    if (nq <= 0) return;
    int mef = 0, mk = 0;
    for (int64_t i = 0; i < nq; ++i) {
      if (k[i] < 0 || ef[i] < 0) fail(NMSAMD_EINVAL, "k and ef must be non-negative");
      mef = std::max(mef, ef[i]);
      mk = std::max(mk, k[i]);
    }
    if (mk > max_k) fail(NMSAMD_EINVAL, "max_k is smaller than a query's k");
    if (max_k <= 0) {
      std::fill(counts, counts + nq, 0);
      return;
    }
    hipStream_t s = h->stream;
    h->d_q.upload(q, (size_t)nq * dim, s);
    h->d_k.upload(k, (size_t)nq, s);
    h->d_ef.upload(ef, (size_t)nq, s);
    // longest-first order: waves pull the largest max(ef, k) queries first so
    // the persistent grid's tail is short
    std::vector<int32_t> order((size_t)nq);
    for (int64_t i = 0; i < nq; ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
      return std::max(ef[x], k[x]) > std::max(ef[y], k[y]);
    });
    h->d_order.upload(order.data(), order.size(), s);
    h->d_ids.ensure((size_t)nq * max_k);
    h->d_dists.ensure((size_t)nq * max_k);
    h->d_counts.ensure((size_t)nq);
    h->order_dev = h->d_order.p;
    run_search(h, h->d_q.p, nq, 0, 0, h->d_k.p, h->d_ef.p, max_k, mef, h->d_ids.p, h->d_dists.p,
               h->d_counts.p, s);
    h->order_dev = nullptr;
    hip_check(hipMemcpyAsync(ids, h->d_ids.p, sizeof(int32_t) * nq * max_k, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(dists, h->d_dists.p, sizeof(float) * nq * max_k, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(counts, h->d_counts.p, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "search");
    finish_stats(h);
  });
}

int64_t nmsamd_size(const nmsamd_index* h) { return h ? (int64_t)h->data_objs.size() : 0; }

int nmsamd_index_info(const nmsamd_index* h, int64_t* o) {
  return guarded([&] {
    if (!h || !o) fail(NMSAMD_EINVAL, "null argument");
    o[0] = h->n; o[1] = h->dim; o[2] = (int64_t)h->maxM; o[3] = (int64_t)h->maxM0;
    o[4] = h->maxlevel; o[5] = h->enter; o[6] = h->func; o[7] = (int64_t)h->ef;
  });
}

int nmsamd_set_precision(nmsamd_index* h, int storage, int rerank_mult) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (storage < kF32 || storage > kFP8) fail(NMSAMD_EINVAL, "storage must be 0 (fp32), 1 (bf16) or 2 (fp8)");
    if (rerank_mult < 0) fail(NMSAMD_EINVAL, "rerank_mult must be >= 0");
    h->storage = storage;
    h->rerank_mult = rerank_mult;
  });
}

int nmsamd_set_beam(nmsamd_index* h, int beam) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (beam < 0 || beam > 64) fail(NMSAMD_EINVAL, "beam must be 0 (exact V1Merge) or 1..64");
    h->beam = beam;
  });
}

int nmsamd_set_beam_gather(nmsamd_index* h, int gather) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (gather < 0 || gather > 2)
      fail(NMSAMD_EINVAL, "beam gather must be 0 (by LDS footprint), 1 (plain) or 2 (wide)");
    h->beam_gather = gather;
  });
}

int nmsamd_set_visited_marks(nmsamd_index* h, int marks) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (marks < 0 || marks > 2)
      fail(NMSAMD_EINVAL, "visited marks must be 0 (by size), 1 (epoch bytes) or 2 (bits)");
    h->visited_marks = marks;
  });
}

int nmsamd_set_screen(nmsamd_index* h, int mode) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (mode < 0 || mode > 2)
      fail(NMSAMD_EINVAL, "screen must be 0 (off), 1 (BF16 screen) or 2 (screen and verify)");
    h->screen = mode;
  });
}

int nmsamd_last_screen_stats(const nmsamd_index* hc, uint64_t* out2) {
  return guarded([&] {
    if (!hc || !out2) fail(NMSAMD_EINVAL, "null argument");
    auto* h = const_cast<nmsamd_index*>(hc);
    if (h->ev1 && hipEventQuery(h->ev1) == hipSuccess) finish_stats(h);
    out2[0] = h->last_stats.screened;
    out2[1] = h->last_stats.screen_pass;
  });
}

int nmsamd_set_beam_waves(nmsamd_index* h, int waves) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    if (waves < 0 || (waves > 2 && waves != 8))
      fail(NMSAMD_EINVAL, "beam waves must be 0 (by LDS footprint), 1, 2 or 8 (cooperative)");
    h->beam_waves = waves;
  });
}

int nmsamd_last_stats(const nmsamd_index* hc, uint64_t* o, double* ms) {
  return guarded([&] {
    if (!hc) fail(NMSAMD_EINVAL, "null handle");
    // the launch may have been asynchronous (device entry point): read the
    // events and counters now, after the caller synchronised
    auto* h = const_cast<nmsamd_index*>(hc);
    if (h->ev1 && hipEventQuery(h->ev1) == hipSuccess) finish_stats(h);
    if (o) {
      o[0] = h->last_stats.expansions;
      o[1] = h->last_stats.dist_level0;
      o[2] = h->last_stats.dist_upper;
      o[3] = h->last_stats.list_ints_read;
      o[4] = h->last_stats.tie_paths;
    }
    if (ms) *ms = h->last_ms;
  });
}

int nmsamd_last_query_us(const nmsamd_index* h, float* out_us, int64_t nq) {
  return guarded([&] {
    if (!h || !out_us) fail(NMSAMD_EINVAL, "null argument");
    if (!h->collect || !h->d_qticks.p) fail(NMSAMD_EINVAL, "statistics collection is off");
    if (nq > h->last_nq) fail(NMSAMD_EINVAL, "more queries requested than the last search held");
    std::vector<uint32_t> t((size_t)nq);
    hip_check(hipMemcpy(t.data(), h->d_qticks.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost),
              "query times");
    for (int64_t i = 0; i < nq; ++i) out_us[i] = (float)(t[i] * 0.01);  // 100 MHz clock
  });
}

int nmsamd_last_launch(const nmsamd_index* h, int64_t* out5) {
  return guarded([&] {
    if (!h || !out5) fail(NMSAMD_EINVAL, "null argument");
    for (int i = 0; i < 5; ++i) out5[i] = h->last_launch[i];
  });
}

int nmsamd_last_phase_us(const nmsamd_index* hc, double* out8) {
  return guarded([&] {
    if (!hc || !out8) fail(NMSAMD_EINVAL, "null argument");
    auto* h = const_cast<nmsamd_index*>(hc);  // device entry point: read back now (see last_stats)
    if (h->ev1 && hipEventQuery(h->ev1) == hipSuccess) finish_stats(h);
    for (int p = 0; p < 8; ++p) out8[p] = h->last_stats.phase_ticks[p] * 0.01;  // 100 MHz
  });
}

int nmsamd_collect_stats(nmsamd_index* h, int on) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    h->collect = on != 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Continuous-batching server (no reference counterpart; the reference serves
// through query_server/, whose knnQueryBatch runs ParallelFor over Search,
// QueryService_server.cpp:411-466).  A server owns `lanes` independent
// launch slots over one loaded index: each lane has its own stream, device and
// pinned host buffers and its own range of visited-mark slots, so batches
// submitted to different lanes run concurrently on the device.  A caller
// admits arriving requests into a batch whenever a lane is free; per-query k
// and ef ride in the batch (the uneven persistent kernel).  Optionally the
// fill + search launch of a lane is captured once as a hipGraph and replayed
// (the query count is read on the device, so the graph's arguments stay fixed).
namespace {

struct ServerLane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr, t0 = nullptr;
  DevBuf<float> q, dists;
  DevBuf<int32_t> k, ef, ids, counts, work;
  DevBuf<int64_t> nq_dev;
  float* hq = nullptr;  // pinned staging
  int32_t *hk = nullptr, *hef = nullptr, *hids = nullptr, *hcounts = nullptr;
  float* hdists = nullptr;
  int64_t* hnq = nullptr;
  int32_t* herr = nullptr;  // the batch's error word (SearchArgs::err)
  int64_t nq = 0;
  bool busy = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t slot_base = 0;
  int blocks = 0;
};

}  // namespace

struct nmsamd_server {
  nmsamd_index* h = nullptr;
  int max_queries = 0, max_ef = 0, max_k = 0, cap = 0, ent = 0, lds = 0, wpb = 0, wide = 0;
  // the search form and device index as they were at open: the LDS layout,
  // the captured graphs and every later launch follow these, whatever the
  // index's setters change afterwards
  int storage = kF32, rerank = 0, beam = 0;
  int pair = 0;  // two waves per query (kModeBeamPair), from the LDS footprint
  int defer = 0;  // ... with the deferred rebuild (a second items buffer)
  int vbits = 0;  // visited marks as bits
  int coop = 0;   // cooperative kernel (one query per block of kCoopWaves waves)
  int coop2 = 0;  // two cooperative blocks per CU
  int screen = 0;  // BF16 screen of the cooperative FP32 l2 search (SearchArgs::screen)
  DevIndex ix{};
  int slots_per_lane = 0;
  bool use_graph = false;
  DevBuf<uint8_t> visited, epoch;
  std::deque<ServerLane> lanes;  // lanes never move once created
  ~nmsamd_server() {
    for (auto& L : lanes) {
      if (L.exec) (void)hipGraphExecDestroy(L.exec);
      if (L.graph) (void)hipGraphDestroy(L.graph);
      if (L.done) (void)hipEventDestroy(L.done);
      if (L.t0) (void)hipEventDestroy(L.t0);
      if (L.stream) (void)hipStreamDestroy(L.stream);
      for (void* p : {(void*)L.hq, (void*)L.hk, (void*)L.hef, (void*)L.hids, (void*)L.hcounts,
                      (void*)L.hdists, (void*)L.hnq, (void*)L.herr})
        if (p) (void)hipHostFree(p);
    }
  }
};

namespace {

SearchArgs server_args(nmsamd_server* s, ServerLane& L) {
  SearchArgs a{};
  a.ix = s->ix;
  a.queries = L.q.p;
  a.nq = 0;
  a.nq_dev = L.nq_dev.p;
  a.k_of = L.k.p;
  a.ef_of = L.ef.p;
  a.cap = s->cap;
  a.out_stride = s->max_k;
  a.out_ids = L.ids.p;
  a.out_dists = L.dists.p;
  a.out_counts = L.counts.p;
  a.work = L.work.p;
  a.err = L.work.p + 1;
  a.storage = s->storage;
  a.rerank = s->rerank;
  a.beam = s->beam;
  a.ent = s->ent;
  a.algo = kAlgoV1Merge;
  a.slot_base = L.slot_base;
  a.defer = s->defer;
  a.vbits = s->vbits;
  a.screen = s->screen;
  return a;
}

void lane_launch(nmsamd_server* s, ServerLane& L) {
  SearchArgs a = server_args(s, L);
  hip_check(launch_fill_i32(L.work.p, 0, 2, L.stream), "fill");
  // the server runs V1Merge, so beam batches always take the beam-only kernels
  const int mode = a.beam > 0 ? (s->coop ? (s->coop2 ? kModeBeamCoop2 : kModeBeamCoop)
                                          : (s->pair ? kModeBeamPair : (s->wide ? kModeBeamWide : kModeBeam)))
                              : kModeAll;
  hip_check(launch_search(a, L.blocks, s->wpb, s->lds, mode, L.stream), "search kernel launch");
}

}  // namespace

extern "C" {

int nmsamd_server_open(nmsamd_index* h, int lanes, int max_queries, int max_ef, int max_k,
                       int use_graph, nmsamd_server** out) {
  return guarded([&] {
    if (!h || !out) fail(NMSAMD_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(h->mu);  // against a concurrent load or build
    require_index(h);
    if (lanes < 1 || lanes > 16 || max_queries < 1 || max_ef < 1 || max_k < 1)
      fail(NMSAMD_EINVAL, "lanes in 1..16 and positive max_queries / max_ef / max_k required");
    if (max_ef >= 1000 && h->algo != kV1Merge)
      fail(NMSAMD_EINVAL, "the server runs V1Merge: keep ef below 1000 or set algoType=v1merge");
    if (h->regular && h->storage != kF32)
      fail(NMSAMD_EINVAL, "reduced-precision traversal needs an optimized index (a regular one is "
                          "searched with the space's own FP32 distance)");
    auto s = std::make_unique<nmsamd_server>();
    s->h = h;
    s->max_queries = max_queries;
    s->max_ef = max_ef;
    s->max_k = max_k;
    s->cap = std::max(max_ef, max_k);
    if (h->storage != kF32 && h->rerank_mult > 0)
      s->cap = std::max(s->cap, std::max(max_ef, max_k) +
                                    std::min(std::max(max_ef, max_k), h->rerank_mult * max_k));
    s->ent = step_entries(h);
    s->coop = use_coop(h) ? 1 : 0;
    s->screen = s->coop ? screen_mode(h) : 0;
    s->lds = lds_wave_bytes(s->cap, h->dstride, s->ent, h->beam > 0);
    if ((size_t)s->lds > kLdsPerCu) fail(NMSAMD_EINVAL, "max_ef / max_k exceed the LDS of one wave");
    ensure_lowp(h, s->screen != 0);  // finished on return: the lanes' streams read the copy at once
    s->storage = h->storage;
    s->rerank = (h->storage != kF32 && h->rerank_mult > 0) ? h->rerank_mult : 0;
    s->beam = h->beam;
    s->vbits = (s->coop || use_bit_marks(h, max_ef, h->beam > 0)) ? 1 : 0;
    // the batch entry points' launch shape: two waves per query where the
    // footprint limits a CU to three queries (one query per block, LDS
    // padded so no fourth block fits), else one, wide where the footprint
    // holds a CU to four waves
    // (the cooperative kernel: one block of kCoopWaves waves per CU)
    const size_t fit = kLdsPerCu / s->lds;
    s->pair = (h->beam > 0 && !s->coop && h->beam_waves != 1 && (h->beam_waves == 2 || fit <= 3))
                  ? 1 : 0;
    if (s->pair) {
      bool defer = false;
      s->lds = pair_lds(s->lds, s->ent, &defer);
      s->defer = defer ? 1 : 0;
    }
    s->wpb = (s->pair || s->coop) ? 1
                                  : (int)std::max<size_t>(1, std::min<size_t>(kWavesPerBlock, fit));
    s->coop2 = (s->coop && coop_two_per_cu(h, fit)) ? 1 : 0;
    const int per_cu = s->coop ? (s->coop2 ? 2 : 1) : (int)std::max<size_t>(
        1, std::min<size_t>(s->pair ? 3 : 8, kLdsPerCu / ((size_t)s->lds * s->wpb)));
    s->wide = (h->beam > 0 && !s->pair && !s->coop && s->wpb * per_cu <= 4) ? 1 : 0;
    // the device's resident waves at this footprint, shared by the lanes
    const int64_t total_blocks = (int64_t)h->cu_count * per_cu;
    // cooperative lanes each launch one block per CU: a batch spreads over the
    // whole device, and concurrent lanes' blocks take CUs as earlier ones finish
    const int64_t blocks_per_lane =
        s->coop ? total_blocks : std::max<int64_t>(1, total_blocks / lanes);
    s->slots_per_lane = (int)(blocks_per_lane * s->wpb);
    const int64_t slots = (int64_t)s->slots_per_lane * lanes;
    // (bit marks of the cooperative kernel: the (n + 128) / 128 16-byte words a
    // query clears; else n + 1 epoch bytes)
    h->vstride_server = s->coop ? (((int64_t)h->n + 128) / 128 * 16 + 255) / 256 * 256
                                : ((int64_t)h->n + 1 + 255) / 256 * 256;
    s->visited.ensure((size_t)slots * h->vstride_server);
    s->epoch.ensure((size_t)slots);
    hip_check(hipMemset(s->visited.p, 0, (size_t)slots * h->vstride_server), "memset");
    hip_check(hipMemset(s->epoch.p, 0, (size_t)slots), "memset");
    s->ix = dev_index(h);
    s->ix.visited = s->visited.p;
    s->ix.epoch = s->epoch.p;
    s->ix.vstride = h->vstride_server;
    s->use_graph = use_graph != 0;
    for (int i = 0; i < lanes; ++i) s->lanes.emplace_back();
    for (int i = 0; i < lanes; ++i) {
      ServerLane& L = s->lanes[i];
      hip_check(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipEventCreateWithFlags(&L.done, hipEventDisableTiming), "hipEventCreate");
      L.q.ensure((size_t)max_queries * h->dim);
      L.dists.ensure((size_t)max_queries * max_k);
      L.ids.ensure((size_t)max_queries * max_k);
      L.counts.ensure((size_t)max_queries);
      L.k.ensure((size_t)max_queries);
      L.ef.ensure((size_t)max_queries);
      L.work.ensure(2);
      L.nq_dev.ensure(1);
      hip_check(hipHostMalloc((void**)&L.hq, sizeof(float) * max_queries * h->dim), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hk, sizeof(int32_t) * max_queries), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hef, sizeof(int32_t) * max_queries), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hids, sizeof(int32_t) * max_queries * max_k), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hdists, sizeof(float) * max_queries * max_k), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hcounts, sizeof(int32_t) * max_queries), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.hnq, sizeof(int64_t)), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&L.herr, sizeof(int32_t)), "hipHostMalloc");
      L.slot_base = (int64_t)i * s->slots_per_lane;
      L.blocks = (int)std::max<int64_t>(1, s->slots_per_lane / s->wpb);
      if (s->use_graph) {
        hip_check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "capture");
        lane_launch(s.get(), L);
        hip_check(hipStreamEndCapture(L.stream, &L.graph), "capture");
        hip_check(hipGraphInstantiate(&L.exec, L.graph, nullptr, nullptr, 0), "graph instantiate");
      }
    }
    ++h->servers;
    *out = s.release();
  });
}

int nmsamd_server_submit(nmsamd_server* s, int lane, const float* q, int64_t nq, const int32_t* k,
                         const int32_t* ef) {
  return guarded([&] {
    if (!s || lane < 0 || lane >= (int)s->lanes.size()) fail(NMSAMD_EINVAL, "bad server lane");
    ServerLane& L = s->lanes[lane];
    if (L.busy) fail(NMSAMD_EINVAL, "lane is busy");
    if (nq < 1 || nq > s->max_queries) fail(NMSAMD_EINVAL, "batch size outside 1..max_queries");
    for (int64_t i = 0; i < nq; ++i) {
      if (k[i] < 1 || k[i] > s->max_k || ef[i] < 0 || ef[i] > s->max_ef)
        fail(NMSAMD_EINVAL, "k or ef outside the server's limits");
    }
    const int dim = s->h->dim;
    std::memcpy(L.hq, q, sizeof(float) * nq * dim);
    std::memcpy(L.hk, k, sizeof(int32_t) * nq);
    std::memcpy(L.hef, ef, sizeof(int32_t) * nq);
    *L.hnq = nq;
    hip_check(hipMemcpyAsync(L.q.p, L.hq, sizeof(float) * nq * dim, hipMemcpyHostToDevice, L.stream), "H2D");
    hip_check(hipMemcpyAsync(L.k.p, L.hk, sizeof(int32_t) * nq, hipMemcpyHostToDevice, L.stream), "H2D");
    hip_check(hipMemcpyAsync(L.ef.p, L.hef, sizeof(int32_t) * nq, hipMemcpyHostToDevice, L.stream), "H2D");
    hip_check(hipMemcpyAsync(L.nq_dev.p, L.hnq, sizeof(int64_t), hipMemcpyHostToDevice, L.stream), "H2D");
    if (L.exec) hip_check(hipGraphLaunch(L.exec, L.stream), "graph launch");
    else lane_launch(s, L);
    const size_t out = (size_t)nq * s->max_k;
    hip_check(hipMemcpyAsync(L.hids, L.ids.p, sizeof(int32_t) * out, hipMemcpyDeviceToHost, L.stream), "D2H");
    hip_check(hipMemcpyAsync(L.hdists, L.dists.p, sizeof(float) * out, hipMemcpyDeviceToHost, L.stream), "D2H");
    hip_check(hipMemcpyAsync(L.hcounts, L.counts.p, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, L.stream), "D2H");
    hip_check(hipMemcpyAsync(L.herr, L.work.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, L.stream), "D2H");
    hip_check(hipEventRecord(L.done, L.stream), "hipEventRecord");
    L.nq = nq;
    L.busy = true;
  });
}

int nmsamd_server_poll(nmsamd_server* s, int lane) {
  if (!s || lane < 0 || lane >= (int)s->lanes.size()) {
    g_err = "bad server lane";
    return NMSAMD_EINVAL;
  }
  ServerLane& L = s->lanes[lane];
  if (!L.busy) return 1;
  const hipError_t e = hipEventQuery(L.done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  g_err = std::string("server lane: ") + hipGetErrorString(e);
  return NMSAMD_EDEVICE;
}

int nmsamd_server_fetch(nmsamd_server* s, int lane, int32_t* ids, float* dists, int32_t* counts) {
  return guarded([&] {
    if (!s || lane < 0 || lane >= (int)s->lanes.size()) fail(NMSAMD_EINVAL, "bad server lane");
    ServerLane& L = s->lanes[lane];
    if (!L.busy) fail(NMSAMD_EINVAL, "lane has no batch");
    hip_check(hipEventSynchronize(L.done), "server lane");
    L.busy = false;
    if (*L.herr & 2)
      fail(NMSAMD_ERUNTIME, "internal error: corrupted search state (a node id outside the index, or an item expanded twice)");
    if (*L.herr & 4)
      fail(NMSAMD_ERUNTIME, "internal error: the BF16 screen's bound failed against an FP32 distance");
    const size_t out = (size_t)L.nq * s->max_k;
    std::memcpy(ids, L.hids, sizeof(int32_t) * out);
    std::memcpy(dists, L.hdists, sizeof(float) * out);
    std::memcpy(counts, L.hcounts, sizeof(int32_t) * L.nq);
  });
}

int nmsamd_server_info(const nmsamd_server* s, int64_t* o) {
  return guarded([&] {
    if (!s || !o) fail(NMSAMD_EINVAL, "null argument");
    o[0] = (int64_t)s->lanes.size(); o[1] = s->max_queries; o[2] = s->max_ef; o[3] = s->max_k;
    o[4] = s->slots_per_lane; o[5] = s->wpb; o[6] = s->use_graph ? 1 : 0;
    o[7] = s->coop ? (s->coop2 ? 4 : 3) : (s->pair ? 2 : s->wide);
  });
}

int nmsamd_server_close(nmsamd_server* s) {
  return guarded([&] {
    if (!s) return;
    for (auto& L : s->lanes)
      if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (s->h) {
      std::lock_guard<std::mutex> g(s->h->mu);
      --s->h->servers;
    }
    delete s;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Exact k-NN by brute force (the device seq_search,
// similarity_search/src/method/seqsearch.cc:143-177): a matrix-core screen
// keeps kc candidates per query, their FP32 distances are recomputed in the
// reference's summation order and ranked by (distance, internal id), and a
// query whose certificate fails (the screen's error could hide an answer) is
// searched again by an exact scan of every point.
namespace nmsamd {
hipError_t launch_brute_norms(const float* vec, float* out, int64_t n, int dstride,
                              hipStream_t stream);
hipError_t launch_brute_screen(const float* vec, const float* xnorm, const float* queries,
                               int64_t n, int64_t nq, int dstride, int func, int kc, int nchunks,
                               float* part_d, int32_t* part_i, float* cand_d, int32_t* cand_i,
                               hipStream_t stream);
hipError_t launch_brute_merge(const float* part_d, const int32_t* part_i, int nlists, int k,
                              int64_t rows, const int32_t* row_of, const int32_t* ext_id,
                              int32_t* ids, float* dists, int32_t* counts, hipStream_t stream);
hipError_t launch_brute_rescore(const DevIndex& ix, const float* queries, int64_t nq, int kc,
                                const float* cand_d, const int32_t* cand_i, const float* eps,
                                int k, int func, int sqrt_out, int32_t* ids, float* dists,
                                int32_t* counts, int32_t* flagged, int* nflag,
                                hipStream_t stream);
hipError_t launch_brute_scan(const DevIndex& ix, const float* queries, const int32_t* rows,
                             const int* nrows, int nrows_max, int k, int func, int nchunks,
                             float* part_d, int32_t* part_i, hipStream_t stream);
}  // namespace nmsamd

namespace {

// Facts about the stored vectors the screen's error bound needs: the largest
// norm, and whether every coordinate is an integer (and its largest magnitude).
void ensure_vector_bounds(nmsamd_index* h) {
  if (h->vec_bounds_known) return;
  double xmax2 = 0, amax = 0;
  bool ints = true;
  for (int64_t i = 0; i < (int64_t)h->n; ++i) {
    const float* v = reinterpret_cast<const float*>(h->block.data() + i * h->elem + 16);
    double s = 0;
    for (int j = 0; j < h->dim; ++j) {
      const double t = v[j];
      s += t * t;
      amax = std::max(amax, std::fabs(t));
      ints = ints && std::nearbyint(t) == t;
    }
    xmax2 = std::max(xmax2, s);
  }
  h->vec_norm_max = std::sqrt(xmax2);
  h->vec_int_absmax = ints ? amax : -1.0;
  h->vec_bounds_known = true;
}

// mode 0: the index's own search distance (what knnQueryBatch reports);
// mode 1: the space's distance as seq_search computes it (l2 with the square
// root in the SSE order, CosineSimilarity, -ScalarProductSIMD)
void run_bruteforce(nmsamd_index* h, const float* q, int64_t nq, int64_t dim, int32_t k,
                    int32_t* ids, float* dists, int32_t* counts, int mode) {
  require_index(h);
  if (dim != h->dim) fail(NMSAMD_EINVAL, "query dimension does not match the index");
  if (h->func == kL1 || h->func == kLInf || h->regular)
    fail(NMSAMD_EINVAL, "brute force on matrix cores covers optimized l2, cosinesimil and "
                        "negdotprod indices");
  if (h->dstride > 128) fail(NMSAMD_EINVAL, "brute force supports dimensions up to 128");
  if (k < 1 || k > 128) fail(NMSAMD_EINVAL, "brute force k must be in 1..128");
  if (nq <= 0) return;
  const bool l2 = h->func == kL2Sqr16 || h->func == kL2Sqr;
  const bool cos = h->func == kNormCos;
  const int exact_func = mode == 0 ? h->func : (l2 ? kSpaceL2 : (cos ? kSpaceCos : kSpaceNegDot));
  hipStream_t s = h->stream;
  ensure_vector_bounds(h);
  if (!h->d_xnorm.p) {
    h->d_xnorm.ensure((size_t)std::max<int64_t>(h->n, 1));
    hip_check(launch_brute_norms(h->d_vec.p, h->d_xnorm.p, h->n, h->dstride, s), "norms");
  }
  // queries padded to the index row stride, normalised for cosine as the
  // index vectors are (hnsw.h:486-497); the screen's error bound per query:
  // an FP32 dot product of d terms is within (d + 2) u sum |q_i x_i| of the
  // exact one, the norms likewise, so |screen - exact| <= 4 (d + 8) u S with
  // S = (|x|max + |q|)^2 for l2 and |x|max |q| for the dot products; both the
  // screen and the reference-order FP32 value are within that of the exact
  // value.  Integer coordinates whose every partial sum stays below 2^24 make
  // the screen exact (bound 0).
  std::vector<float> qp((size_t)nq * h->dstride, 0.f);
  std::vector<float> eps((size_t)nq, 0.f);
  const double u = std::ldexp(1.0, -24);
  for (int64_t i = 0; i < nq; ++i) {
    float* row = &qp[i * h->dstride];
    std::memcpy(row, q + i * dim, sizeof(float) * dim);
    if (cos) {
      float sum = 0.f;
      for (int64_t j = 0; j < dim; ++j) sum = sum + row[j] * row[j];
      if (sum != 0.0f) {
        const float inv = 1.0f / std::sqrt(sum);
        for (int64_t j = 0; j < dim; ++j) row[j] = row[j] * inv;
      }
    }
    double qn2 = 0, qa = 0;
    bool qint = true;
    for (int64_t j = 0; j < dim; ++j) {
      qn2 += (double)row[j] * row[j];
      qa = std::max(qa, (double)std::fabs(row[j]));
      qint = qint && std::nearbyint(row[j]) == row[j];
    }
    const double qn = std::sqrt(qn2), X = h->vec_int_absmax;
    const double lim = std::ldexp(1.0, 24);
    const bool exact = !cos && X >= 0 && qint &&
                       (l2 ? dim * (X * X + qa * qa) < lim : dim * X * qa < lim);
    const double S = l2 ? (h->vec_norm_max + qn) * (h->vec_norm_max + qn) : h->vec_norm_max * qn + 1.0;
    eps[i] = exact ? 0.f : (float)(4.0 * (dim + 8) * u * S);
  }
  const int kc = (int)std::min<int64_t>(std::max<int64_t>(h->n, 1),
                                        std::min(256, std::max(k + 32, 2 * k)));
  const int64_t qblocks = (nq + 31) / 32;
  int nchunks = (int)std::max<int64_t>(1, std::min<int64_t>(16, (2048 + qblocks - 1) / qblocks));
  while (nchunks > 1 && (int64_t)h->n / nchunks < 64) --nchunks;
  h->d_bq.upload(qp.data(), qp.size(), s);
  h->d_beps.upload(eps.data(), eps.size(), s);
  h->d_bpd.ensure((size_t)nq * nchunks * 2 * kc);
  h->d_bpi.ensure((size_t)nq * nchunks * 2 * kc);
  h->d_bcd.ensure((size_t)nq * kc);
  h->d_bci.ensure((size_t)nq * kc);
  h->d_ids.ensure((size_t)nq * k);
  h->d_dists.ensure((size_t)nq * k);
  h->d_counts.ensure((size_t)nq);
  h->d_bflag.ensure((size_t)nq + 1);
  follow_last_launch(h, s);
  hip_check(launch_fill_i32(h->d_bflag.p + nq, 0, 1, s), "fill");
  hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
  hip_check(launch_brute_screen(h->d_vec.p, h->d_xnorm.p, h->d_bq.p, h->n, nq, h->dstride,
                                h->func, kc, nchunks, h->d_bpd.p, h->d_bpi.p, h->d_bcd.p,
                                h->d_bci.p, s),
            "brute force screen");
  DevIndex ix = dev_index(h);
  hip_check(launch_brute_rescore(ix, h->d_bq.p, nq, kc, h->d_bcd.p, h->d_bci.p, h->d_beps.p, k,
                                 exact_func, mode == 1 && l2, h->d_ids.p, h->d_dists.p,
                                 h->d_counts.p, h->d_bflag.p, h->d_bflag.p + nq, s),
            "brute force rescore");
  int nflag = 0;
  hip_check(hipMemcpyAsync(&nflag, h->d_bflag.p + nq, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
  hip_check(hipStreamSynchronize(s), "brute force");
  h->last_bf_rescanned = nflag;
  if (nflag > 0) {
    // exact scans of the listed queries, then their chunk lists merged into
    // their rows
    const int sch = (int)std::max<int64_t>(1, std::min<int64_t>(64, h->n / 4096));
    h->d_bpd.ensure((size_t)nflag * sch * k);
    h->d_bpi.ensure((size_t)nflag * sch * k);
    hip_check(launch_brute_scan(ix, h->d_bq.p, h->d_bflag.p, h->d_bflag.p + nq, nflag, k,
                                exact_func, sch, h->d_bpd.p, h->d_bpi.p, s),
              "brute force exact scan");
    hip_check(launch_brute_merge(h->d_bpd.p, h->d_bpi.p, sch, k, nflag, h->d_bflag.p, h->d_ext.p,
                                 h->d_ids.p, h->d_dists.p, h->d_counts.p, s),
              "brute force merge");
  }
  hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
  hip_check(hipMemcpyAsync(ids, h->d_ids.p, sizeof(int32_t) * nq * k, hipMemcpyDeviceToHost, s), "D2H");
  hip_check(hipMemcpyAsync(dists, h->d_dists.p, sizeof(float) * nq * k, hipMemcpyDeviceToHost, s), "D2H");
  hip_check(hipMemcpyAsync(counts, h->d_counts.p, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, s), "D2H");
  hip_check(hipStreamSynchronize(s), "brute force");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
  h->last_ms = ms;
}

}  // namespace

extern "C" int nmsamd_brute_force(nmsamd_index* h, const float* q, int64_t nq, int64_t dim,
                                  int32_t k, int32_t* ids, float* dists, int32_t* counts) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    run_bruteforce(h, q, nq, dim, k, ids, dists, counts, 0);
  });
}

extern "C" int nmsamd_seq_search(nmsamd_index* h, const float* q, int64_t nq, int64_t dim,
                                 int32_t k, int32_t* ids, float* dists, int32_t* counts) {
  return guarded([&] {
    if (!h) fail(NMSAMD_EINVAL, "null handle");
    std::lock_guard<std::mutex> g(h->mu);
    run_bruteforce(h, q, nq, dim, k, ids, dists, counts, 1);
  });
}

extern "C" int nmsamd_last_bruteforce_rescans(const nmsamd_index* h, int64_t* out) {
  return guarded([&] {
    if (!h || !out) fail(NMSAMD_EINVAL, "null argument");
    *out = h->last_bf_rescanned;
  });
}

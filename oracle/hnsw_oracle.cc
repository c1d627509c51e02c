// ORACLE -- test infrastructure only (see hnsw_oracle.h for the reference map).
// A CPU restatement of nmslib's optimized-HNSW query path.  It is never linked
// by the product; tests and the CPU-baseline leg of bench.py load it.
#include "hnsw_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------- distances
// AVX summation order of hnsw_distfunc_opt_impl_inline.h: eight lane
// accumulators fed by 16-float chunks (lanes j take elements j and j+8 of each
// chunk, mul rounded before add), then either an 8-way left-to-right lane sum
// (L2Sqr16Ext :42-70) or lo+hi folding to 4 lanes, 4-wide steps, a 4-way sum
// and a scalar tail (L2SqrExt :72-122, ScalarProduct :124-173).
// Built with -ffp-contract=off so no FMA contraction happens.
float l2sqr16(const float* a, const float* b, int64_t d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t c = 0; c + 16 <= d; c += 16) {
    for (int j = 0; j < 8; ++j) { float t = a[c + j] - b[c + j]; acc[j] = acc[j] + t * t; }
    for (int j = 0; j < 8; ++j) { float t = a[c + 8 + j] - b[c + 8 + j]; acc[j] = acc[j] + t * t; }
  }
  float s = acc[0];
  for (int j = 1; j < 8; ++j) s = s + acc[j];
  return s;
}

template <bool kDot>
float ext_kernel(const float* a, const float* b, int64_t d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t c = 0;
  for (; c + 16 <= d; c += 16) {
    for (int j = 0; j < 8; ++j) {
      float t = kDot ? a[c + j] * b[c + j] : (a[c + j] - b[c + j]) * (a[c + j] - b[c + j]);
      acc[j] = acc[j] + t;
    }
    for (int j = 0; j < 8; ++j) {
      float t = kDot ? a[c + 8 + j] * b[c + 8 + j]
                     : (a[c + 8 + j] - b[c + 8 + j]) * (a[c + 8 + j] - b[c + 8 + j]);
      acc[j] = acc[j] + t;
    }
  }
  float s4[4];
  for (int j = 0; j < 4; ++j) s4[j] = acc[j] + acc[4 + j];
  for (; c + 4 <= d; c += 4) {
    for (int j = 0; j < 4; ++j) {
      float t = kDot ? a[c + j] * b[c + j] : (a[c + j] - b[c + j]) * (a[c + j] - b[c + j]);
      s4[j] = s4[j] + t;
    }
  }
  float s = ((s4[0] + s4[1]) + s4[2]) + s4[3];
  for (; c < d; ++c) {
    float t = kDot ? a[c] * b[c] : (a[c] - b[c]) * (a[c] - b[c]);
    s = s + t;
  }
  return s;
}

// L1NormSIMD (distcomp_lp.cc:191-247): four SSE lanes fed by groups of 4
// floats in order, the lanes summed left to right in float, the tail added in
// double, the result returned as float.  |a - b| is exact in either order.
float l1_kernel(const float* a, const float* b, int64_t d) {
  float acc[4] = {0, 0, 0, 0};
  int64_t c = 0;
  for (; c + 4 <= d; c += 4)
    for (int j = 0; j < 4; ++j) acc[j] = acc[j] + std::fabs(a[c + j] - b[c + j]);
  double res = ((acc[0] + acc[1]) + acc[2]) + acc[3];
  for (; c < d; ++c) res += std::fabs(a[c] - b[c]);
  return (float)res;
}

// LInfNormSIMD (distcomp_lp.cc:78-139): the largest |a - b|; a maximum does
// not depend on the order of its operands.
float linf_kernel(const float* a, const float* b, int64_t d) {
  float res = 0.0f;
  for (int64_t c = 0; c < d; ++c) res = std::max(res, std::fabs(a[c] - b[c]));
  return res;
}

float dist_fn(int func, const float* q, const float* x, int64_t d) {
  switch (func) {
    case ORA_L1: return l1_kernel(q, x, d);
    case ORA_LINF: return linf_kernel(q, x, d);
    case ORA_L2SQR16: return l2sqr16(q, x, d);
    case ORA_L2SQR: return ext_kernel<false>(q, x, d);
    case ORA_NORMCOS: {
      float dot = ext_kernel<true>(q, x, d);
      return std::max(0.0f, 1.0f - std::max(-1.0f, std::min(1.0f, dot)));
    }
    case ORA_NEGDOT: return -ext_kernel<true>(q, x, d);
    default: throw std::runtime_error("Unknown distance function code: " + std::to_string(func));
  }
}

void normalize(float* v, int64_t d) {
  float sum = 0;
  for (int64_t i = 0; i < d; ++i) sum = sum + v[i] * v[i];
  if (sum != 0.0f) {
    float inv = 1.0f / std::sqrt(sum);
    for (int64_t i = 0; i < d; ++i) v[i] = v[i] * inv;
  }
}

// ---------------------------------------------------------------- index
struct Index {
  uint32_t n = 0;
  uint64_t elem = 0, off_l0 = 0, off_data = 0, maxM = 0, maxM0 = 0, method = 0;
  int32_t maxlevel = 0, func = 0;
  uint32_t enter = 0;
  int64_t dim = 0;
  std::vector<char> block;               // n * elem bytes, as in the file
  std::vector<std::vector<int32_t>> upper;  // per node, level*(maxM+1) ints
  const float* vec(int64_t i) const {
    return reinterpret_cast<const float*>(block.data() + i * elem + off_data + 16);
  }
  const int32_t* l0(int64_t i) const {
    return reinterpret_cast<const int32_t*>(block.data() + i * elem + off_l0);
  }
  int32_t ext_id(int64_t i) const {
    return *reinterpret_cast<const int32_t*>(block.data() + i * elem + off_data);
  }
};

template <class T>
void rd(std::ifstream& in, T& v) { in.read(reinterpret_cast<char*>(&v), sizeof(T)); }

Index* load(const char* path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("Cannot open file '") + path + "' for reading");
  in.exceptions(std::ios::badbit | std::ios::failbit);
  auto ix = std::make_unique<Index>();
  uint32_t flag = 0;
  rd(in, flag);
  if (!flag) throw std::runtime_error("regular (non-optimized) HNSW index format is not supported");
  rd(in, ix->n); rd(in, ix->elem); rd(in, ix->off_l0); rd(in, ix->off_data);
  rd(in, ix->maxlevel); rd(in, ix->enter); rd(in, ix->maxM); rd(in, ix->maxM0);
  rd(in, ix->func); rd(in, ix->method);
  if (ix->func < ORA_L2SQR16 || ix->func > ORA_LINF)
    throw std::runtime_error("Unknown distance function code: " + std::to_string(ix->func));
  ix->dim = ((int64_t)ix->off_l0 - 16) / 4;
  ix->block.resize((size_t)ix->n * ix->elem);
  if (!ix->block.empty()) in.read(ix->block.data(), ix->block.size());
  ix->upper.resize(ix->n);
  for (uint32_t i = 0; i < ix->n; ++i) {
    uint32_t sizemass = 0;
    rd(in, sizemass);
    if (sizemass) {
      ix->upper[i].resize(sizemass / 4);
      in.read(reinterpret_cast<char*>(ix->upper[i].data()), sizemass);
    }
  }
  return ix.release();
}

// ---------------------------------------------------------------- sorted array
// Bounded sorted array of sort_arr_bi.h.  Items carry a key, a "used" flag and
// the internal node number.
struct Item {
  float key;
  bool used;
  int32_t data;
};

struct SortedArr {
  std::vector<Item> v;
  size_t n = 0;
  explicit SortedArr(size_t cap) : v(cap) {}
  void push_grow(float key, int32_t data) {  // push_unsorted_grow (:61-67)
    if (n + 1 >= v.size()) v.resize(n + 1);
    v[n] = Item{key, false, data};
    ++n;
  }
  float top_key() const { return v[n - 1].key; }
  // push_or_replace_non_empty_exp (:159-199): append after an equal last key
  // while there is room, otherwise drop when full; else walk down with steps
  // 1,2,4,.. (clamped to the index) to the first key <= new key and take the
  // lower bound inside the last step's window.  Returns the slot (n if dropped).
  size_t insert(float key, int32_t data) {
    size_t cur = n - 1;
    if (v[cur].key <= key) {
      if (n < v.size()) {
        v[n] = Item{key, false, data};
        return n++;
      }
      return n;
    }
    size_t prev = cur, step = 1;
    while (cur > 0 && v[cur].key > key) {
      prev = cur;
      cur -= step;
      step *= 2;
      if (step > cur) step = cur;
    }
    if (cur < prev) {
      size_t lo = cur, hi = prev;  // first index in [cur, prev) with key >= new key
      while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (v[mid].key < key) lo = mid + 1; else hi = mid;
      }
      cur = lo;
    }
    if (n < v.size()) ++n;
    for (size_t i = n - 1; i > cur; --i) v[i] = v[i - 1];
    v[cur] = Item{key, false, data};
    return cur;
  }
  // merge_with_sorted_items (:123-155), used for more than 100 new items
  size_t merge(Item* items, size_t qty) {
    if (!qty) return n;
    if (qty > v.size()) qty = v.size();
    const size_t left = v.size() - n;
    auto less = [](const Item& a, const Item& b) { return a.key < b.key; };
    if (left >= qty) {
      std::copy(items, items + qty, v.begin() + n);
      std::inplace_merge(v.begin(), v.begin() + n, v.begin() + n + qty, less);
      n += qty;
    } else {
      size_t rem = 0;
      while (qty > left + rem && n > rem && items[left + rem].key < v[n - rem - 1].key) ++rem;
      std::copy(items, items + left + rem, v.begin() + (n - rem));
      std::inplace_merge(v.begin(), v.begin() + (n - rem), v.end(), less);
      n = v.size();
    }
    size_t r = 0;
    while (r < n && v[r].used) ++r;
    return r;
  }
};

// Result heap of KNNQuery/KNNQueue: max-heap of (distance, object) where the
// object order is its address; objects are allocated in internal order, so the
// internal node number stands in for the address.
struct ResultHeap {
  size_t k;
  std::priority_queue<std::pair<float, int32_t>> q;
  explicit ResultHeap(size_t kk) : k(kk) {}
  void offer(float d, int32_t node) {  // CheckAndAddToResult + Push
    if (q.size() < k) {
      q.push({d, node});
    } else if (d < q.top().first) {
      q.pop();
      q.push({d, node});
    }
  }
};

struct Visited {
  std::vector<uint8_t> mass;
  uint8_t cur = 255;
  explicit Visited(size_t n) : mass(n + 1, 0) {}
  void next() {
    ++cur;
    if (cur == 0) {
      std::fill(mass.begin(), mass.end(), 0);
      ++cur;
    }
  }
};

struct OldNode {  // EvaluatedMSWNodeInt (hnsw.h:410-431): ordered by distance only
  float d;
  int32_t e;
  bool operator<(const OldNode& o) const { return d < o.d; }
};

// Greedy descent through the upper levels (hnsw_distfunc_opt.cc:168-198):
// rescan the current node's list at each level, moving to any strictly closer
// neighbour mid-scan, until a full scan changes nothing.
void descend(const Index& ix, const float* q, int32_t& cur, float& curdist, ora_stats& st) {
  const int64_t d = ix.dim;
  cur = (int32_t)ix.enter;
  curdist = dist_fn(ix.func, q, ix.vec(cur), d);
  st.dist_upper++;
  for (int lev = ix.maxlevel; lev > 0; --lev) {
    bool changed = true;
    while (changed) {
      changed = false;
      const int32_t* lst = ix.upper[cur].data() + (ix.maxM + 1) * (lev - 1);
      int32_t cnt = lst[0];
      st.list_ints_read += 1 + cnt;
      st.dist_upper += cnt;
      for (int32_t j = 1; j <= cnt; ++j) {
        int32_t t = lst[j];
        float dd = dist_fn(ix.func, q, ix.vec(t), d);
        if (dd < curdist) {
          curdist = dd;
          cur = t;
          changed = true;
        }
      }
    }
  }
}

void search_one(const Index& ix, float* q, size_t k, size_t ef, int algo, Visited& vis,
                ResultHeap& res, ora_stats& st) {
  if (ix.n == 0) return;
  const int64_t d = ix.dim;
  if (ix.func == ORA_NORMCOS) normalize(q, d);
  vis.next();
  const uint8_t V = vis.cur;

  int32_t cur = 0;
  float curdist = 0.f;
  descend(ix, q, cur, curdist, st);

  const bool use_old = algo == ORA_ALGO_OLD || (algo == ORA_ALGO_HYBRID && ef >= 1000);
  if (use_old) {
    // SearchOld (hnsw_distfunc_opt.cc:46-150)
    std::priority_queue<OldNode> cand, best;
    cand.push({-curdist, cur});
    best.push({curdist, cur});
    res.offer(curdist, cur);
    vis.mass[cur] = V;
    while (!cand.empty()) {
      OldNode c = cand.top();
      if (-c.d > best.top().d) break;
      cand.pop();
      const int32_t* lst = ix.l0(c.e);
      int32_t cnt = lst[0];
      st.expansions++;
      st.list_ints_read += 1 + cnt;
      for (int32_t j = 1; j <= cnt; ++j) {
        int32_t t = lst[j];
        if (vis.mass[t] == V) continue;
        vis.mass[t] = V;
        st.dist_level0++;
        float dd = dist_fn(ix.func, q, ix.vec(t), d);
        if (best.top().d > dd || best.size() < ef) {
          cand.push({-dd, t});
          res.offer(dd, t);
          best.push({dd, t});
          if (best.size() > ef) best.pop();
        }
      }
    }
    return;
  }

  // SearchV1Merge (hnsw_distfunc_opt.cc:152-283)
  SortedArr arr(std::max(ef, k));
  arr.push_grow(curdist, cur);
  size_t curr = 0;
  std::vector<Item> buf(1 + std::max(ix.maxM, ix.maxM0));
  vis.mass[cur] = V;
  while (curr < std::min(arr.n, ef)) {
    Item& e = arr.v[curr];
    e.used = true;
    int32_t node = e.data;
    ++curr;
    size_t nb = 0;
    float topk = arr.top_key();
    const int32_t* lst = ix.l0(node);
    int32_t cnt = lst[0];
    st.expansions++;
    st.list_ints_read += 1 + cnt;
    for (int32_t j = 1; j <= cnt; ++j) {
      int32_t t = lst[j];
      if (vis.mass[t] == V) continue;
      vis.mass[t] = V;
      st.dist_level0++;
      float dd = dist_fn(ix.func, q, ix.vec(t), d);
      if (dd < topk || arr.n < ef) {
        if (nb >= buf.size()) throw std::runtime_error("Perhaps a bug: buffer size is not enough");
        buf[nb++] = Item{dd, false, t};
      }
    }
    if (nb) {
      std::sort(buf.begin(), buf.begin() + nb,
                [](const Item& a, const Item& b) { return a.key < b.key; });
      if (nb > 100) {
        size_t pos = arr.merge(buf.data(), nb);
        if (pos < curr) curr = pos;
      } else {
        for (size_t i = 0; i < nb; ++i) {
          size_t pos = arr.insert(buf[i].key, buf[i].data);
          if (pos < curr) curr = pos;
        }
      }
    }
    while (curr < arr.n && arr.v[curr].used) ++curr;
  }
  for (size_t i = 0; i < k && i < arr.n; ++i) res.offer(arr.v[i].key, arr.v[i].data);
}

// Beam mode of the product (an extension with no reference counterpart; this
// model is its checker).  V1Merge's array (capacity max(ef, k), stop once the
// first min(size, ef) items are all expanded) ordered by (distance, internal
// id); each step expands the first `beam` unexpanded items together: every
// neighbour unvisited before the step is taken once, the snapshot rule of
// hnsw_distfunc_opt.cc:240-246 (d < last key, or fewer than ef items) admits
// it, and the admitted items join the array in (distance, id) order.
struct BeamItem {
  float key;
  int32_t id;
  bool used;
  bool operator<(const BeamItem& o) const { return key < o.key || (key == o.key && id < o.id); }
};

// One step of the beam model as the device's two-wave protocol sees it: the
// array size before the step, the picks, the fresh neighbours, the admitted
// items, and for each merge pass of at most 256 items (the device's pass
// size) the position its smallest item takes.
struct BeamStep {
  int32_t n, nsel, ncand, m, p0[2];
};

void search_beam_one(const Index& ix, float* q, size_t k, size_t ef, int beam, Visited& vis,
                     std::vector<BeamItem>& A, ora_stats& st,
                     std::vector<BeamStep>* trace = nullptr) {
  A.clear();
  if (ix.n == 0 || k == 0) return;
  const int64_t d = ix.dim;
  if (ix.func == ORA_NORMCOS) normalize(q, d);
  vis.next();
  const uint8_t V = vis.cur;
  int32_t cur = 0;
  float curdist = 0.f;
  descend(ix, q, cur, curdist, st);
  const size_t C = std::max(ef, k);
  A.push_back({curdist, cur, false});
  vis.mass[cur] = V;
  size_t ce = 0;
  std::vector<int32_t> picks, fresh;
  std::vector<BeamItem> add;
  while (ce < std::min(A.size(), ef)) {
    const size_t lim = std::min(A.size(), ef);
    picks.clear();
    for (size_t p = ce; p < lim && (int)picks.size() < beam; ++p)
      if (!A[p].used) {
        A[p].used = true;
        picks.push_back(A[p].id);
      }
    st.expansions += (int64_t)picks.size();
    const float topk = A.back().key;
    const size_t nsnap = A.size();
    fresh.clear();
    for (int32_t e : picks) {
      const int32_t* lst = ix.l0(e);
      const int32_t cnt = lst[0];
      st.list_ints_read += 1 + cnt;
      for (int32_t j = 1; j <= cnt; ++j) {
        const int32_t t = lst[j];
        if (vis.mass[t] != V && std::find(fresh.begin(), fresh.end(), t) == fresh.end())
          fresh.push_back(t);
      }
    }
    add.clear();
    for (int32_t t : fresh) {
      vis.mass[t] = V;
      const float dd = dist_fn(ix.func, q, ix.vec(t), d);
      st.dist_level0++;
      if (dd < topk || nsnap < ef) add.push_back({dd, t, false});
    }
    std::sort(add.begin(), add.end());
    if (trace) {
      BeamStep bs{(int32_t)A.size(), (int32_t)picks.size(), (int32_t)fresh.size(),
                  (int32_t)add.size(), {-1, -1}};
      // passes of 256 merged one after another: the first item of pass p lands
      // after the A items below it and the earlier passes' items below it
      // (all of them: the passes are in order), truncated at C in between
      size_t n = A.size();
      std::vector<BeamItem> cur(A.begin(), A.end());
      for (size_t b = 0, p = 0; b < add.size() && p < 2; b += 256, ++p) {
        const size_t cnt = std::min<size_t>(256, add.size() - b);
        bs.p0[p] = (int32_t)(std::lower_bound(cur.begin(), cur.end(), add[b]) - cur.begin());
        std::vector<BeamItem> mg(cur.size() + cnt);
        std::merge(cur.begin(), cur.end(), add.begin() + b, add.begin() + b + cnt, mg.begin());
        if (mg.size() > C) mg.resize(C);
        cur.swap(mg);
        n = cur.size();
      }
      (void)n;
      trace->push_back(bs);
    }
    std::vector<BeamItem> merged(A.size() + add.size());
    std::merge(A.begin(), A.end(), add.begin(), add.end(), merged.begin());
    if (merged.size() > C) merged.resize(C);
    A.swap(merged);
    ce = 0;
    while (ce < A.size() && A[ce].used) ++ce;
  }
}

template <class F>
void parallel_for(int64_t n, int threads, F fn) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (threads == 1 || n <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> pool;
  std::exception_ptr err;
  std::mutex m;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      try {
        for (int64_t i; (i = next.fetch_add(1)) < n;) fn(i, t);
      } catch (...) {
        std::lock_guard<std::mutex> g(m);
        err = std::current_exception();
      }
    });
  }
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace

extern "C" {

const char* ora_last_error(void) { return g_err.c_str(); }

void* ora_load(const char* path) {
  try {
    return load(path);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void ora_free(void* h) { delete static_cast<Index*>(h); }

void ora_info(void* h, int64_t* o) {
  auto* ix = static_cast<Index*>(h);
  o[0] = ix->n; o[1] = ix->dim; o[2] = (int64_t)ix->maxM; o[3] = (int64_t)ix->maxM0;
  o[4] = ix->maxlevel; o[5] = ix->enter; o[6] = ix->func; o[7] = (int64_t)ix->elem;
}

const char* ora_level0(void* h) { return static_cast<Index*>(h)->block.data(); }

int32_t ora_level_of(void* h, int64_t i) {
  auto* ix = static_cast<Index*>(h);
  return (int32_t)(ix->upper[i].size() / (ix->maxM + 1));
}

const int32_t* ora_upper_list(void* h, int64_t i, int level) {
  auto* ix = static_cast<Index*>(h);
  return ix->upper[i].data() + (ix->maxM + 1) * (level - 1);
}

int ora_search(void* h, const float* q, int64_t nq, int k, int ef, int algo, int num_threads,
               int32_t* ids, float* dists, int32_t* counts, ora_stats* stats) {
  try {
    const Index& ix = *static_cast<Index*>(h);
    if (k < 0 || ef < 0) throw std::invalid_argument("k and ef must be non-negative");
    int threads = num_threads <= 0 ? (int)std::max(1u, std::thread::hardware_concurrency())
                                   : num_threads;
    std::vector<Visited> vis;
    for (int t = 0; t < threads; ++t) vis.emplace_back(ix.n);
    std::vector<ora_stats> st(threads, ora_stats{0, 0, 0, 0});
    parallel_for(nq, threads, [&](int64_t i, int t) {
      std::vector<float> qv(q + i * ix.dim, q + (i + 1) * ix.dim);
      ResultHeap res((size_t)k);
      if (k > 0) search_one(ix, qv.data(), (size_t)k, (size_t)ef, algo, vis[t], res, st[t]);
      size_t sz = res.q.size();
      counts[i] = (int32_t)sz;
      for (int j = (int)sz; j < k; ++j) {
        ids[i * k + j] = -1;
        dists[i * k + j] = 0.f;
      }
      while (!res.q.empty()) {
        --sz;
        ids[i * k + sz] = ix.ext_id(res.q.top().second);
        dists[i * k + sz] = res.q.top().first;
        res.q.pop();
      }
    });
    if (stats) {
      for (auto& s : st) {
        stats->expansions += s.expansions;
        stats->dist_level0 += s.dist_level0;
        stats->dist_upper += s.dist_upper;
        stats->list_ints_read += s.list_ints_read;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// The beam model's steps for one query (BeamStep records, 6 ints each, up to
// cap steps); returns the number of steps, or -1 on error.
int64_t ora_beam_trace(void* h, const float* q, int k, int ef, int beam, int32_t* out,
                       int64_t cap) {
  try {
    const Index& ix = *static_cast<Index*>(h);
    if (k < 0 || ef < 0 || beam < 1) throw std::invalid_argument("k, ef >= 0 and beam >= 1 required");
    Visited vis(ix.n);
    std::vector<float> qv(q, q + ix.dim);
    std::vector<BeamItem> A;
    ora_stats st{0, 0, 0, 0};
    std::vector<BeamStep> tr;
    search_beam_one(ix, qv.data(), (size_t)k, (size_t)ef, beam, vis, A, st, &tr);
    for (size_t i = 0; i < tr.size() && (int64_t)i < cap; ++i)
      std::memcpy(out + 6 * i, &tr[i], sizeof(BeamStep));
    return (int64_t)tr.size();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int ora_search_beam(void* h, const float* q, int64_t nq, int k, int ef, int beam, int num_threads,
                    int32_t* ids, float* dists, int32_t* counts, ora_stats* stats) {
  try {
    const Index& ix = *static_cast<Index*>(h);
    if (k < 0 || ef < 0 || beam < 1) throw std::invalid_argument("k, ef >= 0 and beam >= 1 required");
    int threads = num_threads <= 0 ? (int)std::max(1u, std::thread::hardware_concurrency())
                                   : num_threads;
    std::vector<Visited> vis;
    for (int t = 0; t < threads; ++t) vis.emplace_back(ix.n);
    std::vector<ora_stats> st(threads, ora_stats{0, 0, 0, 0});
    parallel_for(nq, threads, [&](int64_t i, int t) {
      std::vector<float> qv(q + i * ix.dim, q + (i + 1) * ix.dim);
      std::vector<BeamItem> A;
      search_beam_one(ix, qv.data(), (size_t)k, (size_t)ef, beam, vis[t], A, st[t]);
      const size_t m = std::min(A.size(), (size_t)k);
      counts[i] = (int32_t)m;
      for (int j = 0; j < k; ++j) {
        ids[i * k + j] = j < (int)m ? ix.ext_id(A[j].id) : -1;
        dists[i * k + j] = j < (int)m ? A[j].key : 0.f;
      }
    });
    if (stats) {
      for (auto& s : st) {
        stats->expansions += s.expansions;
        stats->dist_level0 += s.dist_level0;
        stats->dist_upper += s.dist_upper;
        stats->list_ints_read += s.list_ints_read;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

float ora_dist(int func, const float* q, const float* x, int64_t dim) {
  try {
    return dist_fn(func, q, x, dim);
  } catch (const std::exception& e) {
    g_err = e.what();
    return NAN;
  }
}

void ora_normalize(float* v, int64_t dim) { normalize(v, dim); }

int ora_brute_force(const float* data, int64_t n, int64_t dim, const float* q, int64_t nq, int k,
                    int space, int num_threads, int64_t* ids, double* dists) {
  try {
    if (space < 1 || space > 5) throw std::invalid_argument("bad space code");
    std::vector<double> xnorm(space == 2 ? n : 0);
    for (int64_t i = 0; i < (int64_t)xnorm.size(); ++i) {
      double s = 0;
      for (int64_t j = 0; j < dim; ++j) s += (double)data[i * dim + j] * data[i * dim + j];
      xnorm[i] = std::sqrt(s);
    }
    parallel_for(nq, num_threads, [&](int64_t qi, int) {
      const float* qq = q + qi * dim;
      double qn = 0;
      for (int64_t j = 0; j < dim; ++j) qn += (double)qq[j] * qq[j];
      qn = std::sqrt(qn);
      std::vector<std::pair<double, int64_t>> all((size_t)n);
      for (int64_t i = 0; i < n; ++i) {
        const float* x = data + i * dim;
        double s = 0;
        if (space == 4) {
          for (int64_t j = 0; j < dim; ++j) s += std::fabs((double)qq[j] - x[j]);
        } else if (space == 5) {
          for (int64_t j = 0; j < dim; ++j) s = std::max(s, std::fabs((double)qq[j] - x[j]));
        } else if (space == 1) {
          for (int64_t j = 0; j < dim; ++j) { double t = (double)qq[j] - x[j]; s += t * t; }
        } else {
          for (int64_t j = 0; j < dim; ++j) s += (double)qq[j] * x[j];
          if (space == 2) {
            double den = qn * xnorm[i];
            // zero vectors: cosine similarity 0 (distcomp_scalar.cc:83-168 returns 0)
            double cs = den > 0 ? std::max(-1.0, std::min(1.0, s / den)) : 0.0;
            s = 1.0 - cs;
          } else {
            s = -s;
          }
        }
        all[i] = {s, i};
      }
      size_t kk = std::min<size_t>(k, all.size());
      std::partial_sort(all.begin(), all.begin() + kk, all.end());
      for (size_t j = 0; j < (size_t)k; ++j) {
        ids[qi * k + j] = j < kk ? all[j].second : -1;
        dists[qi * k + j] = j < kk ? all[j].first : 0.0;
      }
    });
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"

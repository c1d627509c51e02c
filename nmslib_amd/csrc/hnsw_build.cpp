// HNSW index construction on the host CPU, producing the optimized index the
// device query path loads.  Reference counterparts (similarity_search/):
//   build_hnsw                  Hnsw::CreateIndex          src/method/hnsw.cc:183-470
//   Graph::add                  Hnsw::add                  src/method/hnsw.cc:534-611
//   Graph::search_level         kSearchElementsWithAttemptsLevel hnsw.cc:613-708
//   Graph::select_h1 / _h2 / _h3 getNeighborsByHeuristic1/2/3 include/method/hnsw.h:82-258
//   Graph::add_friend           HnswNode::addFriendlevel   include/method/hnsw.h:260-307
//   random_level                Hnsw::getRandomLevel       include/method/hnsw.h:478-483
//   dist_*                      the spaces' HiddenDistance (space_lp.cc:27-35,
//                               space_scalar.cc:28-68) over distcomp_lp.cc and
//                               distcomp_scalar.cc
//   layout                      CreateIndex's optimized copy, hnsw.cc:362-470
//
// With one build thread the graph equals the reference's exactly, so the saved
// file does too (apart from the unused slots of upper-level lists, which the
// reference leaves uninitialised and this writes as zeros).  That rests on
// three things being the same as the reference's:
//   * level draws: a per-thread std::mt19937 seeded with 0 (the library seed,
//     init.cc:34, python_bindings/nmslib.cc:450) feeding
//     std::uniform_real_distribution<float> (utils.h:96-129).  As there, the
//     calling thread's generator carries on from one build to the next, and a
//     fresh thread starts again from seed 0.
//   * index-time distances bit for bit: four-lane SSE accumulators in the
//     order of the reference's SIMD kernels, separate multiply and add
//     roundings (this file is compiled with -ffp-contract=off).
//   * heap orders: std::priority_queue over the same distance-only
//     comparisons, fed the same sequence of pushes and pops.
// With several threads the insertion order depends on scheduling, as in the
// reference; every node is still inserted by the same procedure.
#include "hnsw_build.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_set>

namespace nmsamd {
namespace {

// ---------------------------------------------------------------- distances
// Four SSE lanes take elements 4j..4j+3 in turn (the reference unrolls the
// same sequence of 4-wide steps by four), the lanes are summed left to right,
// and the scalar tail is added last.

float l2_sqr_sse(const float* a, const float* b, size_t n) {  // L2SqrSIMD, distcomp_lp.cc:304-361
  const size_t n4 = n / 4 * 4;
  __m128 acc = _mm_setzero_ps();
  size_t i = 0;
  for (; i < n4; i += 4) {
    const __m128 d = _mm_sub_ps(_mm_loadu_ps(a + i), _mm_loadu_ps(b + i));
    acc = _mm_add_ps(acc, _mm_mul_ps(d, d));
  }
  alignas(16) float t[4];
  _mm_store_ps(t, acc);
  float r = t[0] + t[1] + t[2] + t[3];
  for (; i < n; ++i) {
    const float d = a[i] - b[i];
    r += d * d;
  }
  return r;
}

float dot_sse(const float* a, const float* b, size_t n) {  // ScalarProductSIMD, distcomp_scalar.cc:194-245
  const size_t n4 = n / 4 * 4;
  __m128 acc = _mm_setzero_ps();
  size_t i = 0;
  for (; i < n4; i += 4) acc = _mm_add_ps(acc, _mm_mul_ps(_mm_loadu_ps(a + i), _mm_loadu_ps(b + i)));
  alignas(16) float t[4];
  _mm_store_ps(t, acc);
  float r = t[0] + t[1] + t[2] + t[3];
  for (; i < n; ++i) r += a[i] * b[i];
  return r;
}

// NormScalarProductSIMD (distcomp_scalar.cc:83-168): the dot product and both
// squared norms in three lane sets; 0 when either norm is below 2 FLT_MIN,
// else sum / sqrt(norm1) / sqrt(norm2) clamped to [-1, 1].
float norm_dot_sse(const float* a, const float* b, size_t n) {
  const size_t n4 = n / 4 * 4;
  __m128 sp = _mm_setzero_ps(), s1 = sp, s2 = sp;
  size_t i = 0;
  for (; i < n4; i += 4) {
    const __m128 x = _mm_loadu_ps(a + i), y = _mm_loadu_ps(b + i);
    sp = _mm_add_ps(sp, _mm_mul_ps(x, y));
    s1 = _mm_add_ps(s1, _mm_mul_ps(x, x));
    s2 = _mm_add_ps(s2, _mm_mul_ps(y, y));
  }
  alignas(16) float tp[4], t1[4], t2[4];
  _mm_store_ps(tp, sp);
  _mm_store_ps(t1, s1);
  _mm_store_ps(t2, s2);
  float sum = tp[0] + tp[1] + tp[2] + tp[3];
  float norm1 = t1[0] + t1[1] + t1[2] + t1[3];
  float norm2 = t2[0] + t2[1] + t2[2] + t2[3];
  for (; i < n; ++i) {
    sum += a[i] * b[i];
    norm1 += a[i] * a[i];
    norm2 += b[i] * b[i];
  }
  const float eps = std::numeric_limits<float>::min() * 2;
  if (norm1 < eps || norm2 < eps) return 0;
  return std::max(-1.0f, std::min(1.0f, sum / std::sqrt(norm1) / std::sqrt(norm2)));
}

// L1NormSIMD (distcomp_lp.cc:191-247): lane sums in float, the lane total
// widened to double, the tail added in double, the result narrowed to float.
float l1_sse(const float* a, const float* b, size_t n) {
  const size_t n4 = n / 4 * 4;
  const __m128 mask = _mm_castsi128_ps(_mm_set1_epi32(0x7fffffff));
  __m128 acc = _mm_setzero_ps();
  size_t i = 0;
  for (; i < n4; i += 4)
    acc = _mm_add_ps(acc, _mm_and_ps(_mm_sub_ps(_mm_loadu_ps(a + i), _mm_loadu_ps(b + i)), mask));
  alignas(16) float t[4];
  _mm_store_ps(t, acc);
  double r = t[0] + t[1] + t[2] + t[3];
  for (; i < n; ++i) r += std::fabs(a[i] - b[i]);
  return (float)r;
}

// LInfNormSIMD (distcomp_lp.cc:78-139): the largest |a_i - b_i|.
float linf_sse(const float* a, const float* b, size_t n) {
  float r = 0;
  for (size_t i = 0; i < n; ++i) r = std::max(r, std::fabs(a[i] - b[i]));
  return r;
}

// NormalizeVect (include/method/hnsw.h:486-497): sequential float sum of
// squares, then a multiply by 1/sqrt(sum); all-zero vectors stay as they are.
void normalize_vect(float* v, size_t n) {
  float s = 0;
  for (size_t i = 0; i < n; ++i) s += v[i] * v[i];
  if (s != 0) {
    s = 1 / std::sqrt(s);
    for (size_t i = 0; i < n; ++i) v[i] *= s;
  }
}

// getRandomLevel: floor(-ln(U) * mult), U ~ uniform_real_distribution<float>
// over the thread's generator (seed 0).
int random_level(double mult) {
  static thread_local std::mt19937 gen(0);
  static thread_local std::uniform_real_distribution<float> unit(0, 1);
  const float r = -std::log(unit(gen)) * mult;
  return (int)r;
}

// ---------------------------------------------------------------- graph
// Heap entries.  `Near` orders a max-heap by distance (its top is the
// farthest: HnswNodeDistCloser); `Far` inverts it (top = nearest:
// HnswNodeDistFarther).  Both compare the distance only.
struct Near {
  float d;
  int32_t id;
  bool operator<(const Near& o) const { return d < o.d; }
};
struct Far {
  float d;
  int32_t id;
  bool operator<(const Far& o) const { return d > o.d; }
};
using NearHeap = std::priority_queue<Near>;
using FarHeap = std::priority_queue<Far>;

struct Node {
  int level = 0;
  std::vector<std::vector<int32_t>> links;  // per level, in list order
  std::mutex mu;
};

// Epoch marks of one build thread (VisitedList, hnsw.h:568-600).
struct Visited {
  std::vector<uint32_t> mark;
  uint32_t epoch = 0;
  explicit Visited(size_t n) : mark(n, 0) {}
  void next() {
    if (++epoch == 0) {
      std::fill(mark.begin(), mark.end(), 0);
      epoch = 1;
    }
  }
};

class Graph {
 public:
  Graph(int space, const std::vector<const float*>& vec, size_t dim, const BuildParams& p)
      : space_(space), vec_(vec), dim_(dim), p_(p), n_(vec.size()) {}

  // IndexTimeDistance(a, b): `a` is the left operand of the space's function.
  float dist(int32_t a, int32_t b) const {
    const float* x = vec_[a];
    const float* y = vec_[b];
    switch (space_) {
      case kBuildL2: return std::sqrt(l2_sqr_sse(x, y, dim_));
      case kBuildCosine: {
        const float v = std::max(0.0f, 1 - norm_dot_sse(x, y, dim_));
        if (std::isnan(v)) throw std::runtime_error("Bug: NAN dist! (SpaceCosineSimilarity)");
        return v;
      }
      case kBuildNegDot: return -dot_sse(x, y, dim_);
      case kBuildL1: return l1_sse(x, y, dim_);
      default: return linf_sse(x, y, dim_);
    }
  }

  // One insertion pass over all points: point 0 first, then 1..N-1 (or
  // N-1..1 for the reverse pass of post-processing).
  void run_pass(bool reverse) {
    nodes_.reset(new Node[n_]);
    const int l0 = random_level(p_.mult);
    init_node(0, l0);
    top_.store(pack(l0, 0));
    std::atomic<size_t> done(1);
    const size_t step = std::max<size_t>(1, n_ / 20);
    auto body = [&](size_t pos, Visited& vis) {
      const int32_t id = (int32_t)(reverse ? n_ - pos : pos);
      add(id, vis);
      if (p_.progress) {  // the reference's ProgressDisplay on stderr, every 5 %
        const size_t d = done.fetch_add(1) + 1;
        if (d % step == 0 || d == n_)
          std::fprintf(stderr, "[createIndex] %s%zu / %zu points inserted\n",
                       reverse ? "reverse pass: " : "", d, n_);
      }
    };
    parallel(1, n_, body);
  }

  // post = 1 or 2: merge each level-0 list with that of the previous graph
  // (CreateIndex, hnsw.cc:268-335).  Returns the new maxM0 for post = 1.
  size_t merge_with(const std::unique_ptr<Node[]>& prev) {
    std::vector<std::vector<int32_t>> rez(n_);
    std::vector<size_t> union_size(n_, 0);
    auto body = [&](size_t id, Visited&) {
      std::unordered_set<size_t> both;
      for (int32_t v : nodes_[id].links[0]) both.insert((size_t)v);
      for (int32_t v : prev[id].links[0]) both.insert((size_t)v);
      union_size[id] = both.size();
      auto& out = rez[id];
      if (p_.post == 2) {
        NearHeap rs;
        for (size_t cur : both) rs.push({dist((int32_t)cur, (int32_t)id), (int32_t)cur});
        if (p_.delaunay == 0) {
          while (rs.size() > p_.maxM0) rs.pop();
        } else if (p_.delaunay == 3) {
          select_h3((int32_t)id, rs, (int)p_.maxM0, 0);
        } else {
          select_h1(rs, (int)p_.maxM0);
        }
        while (!rs.empty()) {
          out.push_back(rs.top().id);
          rs.pop();
        }
        // heuristic 3 reads other lists: as in the reference's loop, a list
        // is replaced as soon as it is made (single-threaded)
        if (p_.delaunay == 3) nodes_[id].links[0].swap(out);
      } else {
        for (size_t cur : both) out.push_back((int32_t)cur);
      }
    };
    parallel(1, n_, body);
    if (!(p_.post == 2 && p_.delaunay == 3))
      for (size_t id = 1; id < n_; ++id) nodes_[id].links[0].swap(rez[id]);
    size_t mx = 0;
    for (size_t id = 1; id < n_; ++id) mx = std::max(mx, union_size[id]);
    return mx;
  }

  std::unique_ptr<Node[]> take_nodes() { return std::move(nodes_); }
  const Node& node(size_t i) const { return nodes_[i]; }
  int maxlevel() const { return (int)(top_.load() >> 32); }
  int32_t enter() const { return (int32_t)(top_.load() & 0xffffffffu); }

 private:
  static uint64_t pack(int level, int32_t id) { return ((uint64_t)(uint32_t)level << 32) | (uint32_t)id; }

  void init_node(int32_t id, int level) {
    std::lock_guard<std::mutex> g(nodes_[id].mu);
    nodes_[id].level = level;
    nodes_[id].links.assign(level + 1, {});
  }

  void copy_links(int32_t id, int level, std::vector<int32_t>& buf) {
    std::lock_guard<std::mutex> g(nodes_[id].mu);
    buf = nodes_[id].links[level];
  }

  template <class F>
  void parallel(size_t begin, size_t end, F&& fn) {
    const size_t threads = std::max<size_t>(1, p_.threads);
    if (threads == 1) {  // the calling thread does the work (and draws the levels)
      Visited vis(n_);
      for (size_t i = begin; i < end; ++i) fn(i, vis);
      return;
    }
    std::atomic<size_t> next(begin);
    std::exception_ptr err;
    std::mutex err_mu;
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; ++t) {
      pool.emplace_back([&] {
        try {
          Visited vis(n_);
          for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= end) break;
            fn(i, vis);
          }
        } catch (...) {
          std::lock_guard<std::mutex> g(err_mu);
          err = std::current_exception();
          next.store(end);
        }
      });
    }
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
  }

  // Hnsw::add: draw a level, descend greedily to it through the levels above,
  // then at each level from there down to 0 search with efConstruction,
  // select neighbours and link both ways.
  void add(int32_t id, Visited& vis) {
    const int level = random_level(p_.mult);
    std::unique_lock<std::mutex> grow;
    if (level > (int)(top_.load() >> 32)) grow = std::unique_lock<std::mutex>(grow_mu_);
    init_node(id, level);
    const uint64_t t = top_.load();
    const int maxlevel = (int)(t >> 32);
    int32_t ep = (int32_t)(t & 0xffffffffu);
    std::vector<int32_t> buf;
    if (level < maxlevel) {
      float cur_d = dist(id, ep);
      int32_t cur = ep;
      for (int L = maxlevel; L > level; --L) {
        bool changed = true;
        while (changed) {
          changed = false;
          copy_links(cur, L, buf);  // the scan continues over this list if `cur` moves
          for (int32_t nb : buf) __builtin_prefetch(vec_[nb]);
          for (int32_t nb : buf) {
            const float d = dist(id, nb);
            if (d < cur_d) {
              cur_d = d;
              cur = nb;
              changed = true;
            }
          }
        }
      }
      ep = cur;
    }
    for (int L = std::min(level, maxlevel); L >= 0; --L) {
      NearHeap rs;
      search_level(id, rs, ep, L, vis, buf);
      if (p_.delaunay == 0) {
        while (rs.size() > p_.M) rs.pop();
      } else if (p_.delaunay == 1) {
        select_h1(rs, (int)p_.M);
      } else if (p_.delaunay == 3) {
        select_h3(id, rs, (int)p_.M, L);
      } else {
        select_h2(rs, (int)p_.M);
      }
      while (!rs.empty()) {  // farthest first; the last (nearest) seeds the next level
        ep = rs.top().id;
        add_friend(rs.top().id, id, L);
        add_friend(id, rs.top().id, L);
        rs.pop();
      }
    }
    const int enter_level = nodes_[(int32_t)(top_.load() & 0xffffffffu)].level;
    if (level > enter_level) top_.store(pack(level, id));
  }

  // Best-first search of one level bounded by efConstruction: candidates
  // leave nearest first, the result heap keeps the efConstruction nearest,
  // and the search stops once the nearest candidate is farther than the
  // farthest result.
  void search_level(int32_t q, NearHeap& rs, int32_t ep, int L, Visited& vis,
                    std::vector<int32_t>& buf) {
    vis.next();
    FarHeap cand;
    const float d0 = dist(q, ep);
    cand.push({d0, ep});
    rs.push({d0, ep});
    vis.mark[ep] = vis.epoch;
    while (!cand.empty()) {
      const Far cur = cand.top();
      if (cur.d > rs.top().d) break;
      copy_links(cur.id, L, buf);
      cand.pop();
      for (int32_t nb : buf) {  // start the vector and mark loads (hnsw.cc:658-662)
        __builtin_prefetch(vec_[nb]);
        __builtin_prefetch(&vis.mark[nb]);
      }
      for (int32_t nb : buf) {
        if (vis.mark[nb] == vis.epoch) continue;
        vis.mark[nb] = vis.epoch;
        const float d = dist(q, nb);
        if (rs.size() < p_.efConstruction || rs.top().d > d) {
          rs.push({d, nb});
          cand.push({d, nb});
          if (rs.size() > p_.efConstruction) rs.pop();
        }
      }
    }
  }

  // Heuristic 2: visit candidates nearest first and keep one unless a kept
  // neighbour is strictly closer to it than the new point is; stop at NN.
  void select_h2(NearHeap& rs, int nn) const {
    if (rs.size() < (size_t)nn) return;
    FarHeap q;
    while (!rs.empty()) {
      q.push({rs.top().d, rs.top().id});
      rs.pop();
    }
    std::vector<Far> keep;
    while (!q.empty()) {
      if (keep.size() >= (size_t)nn) break;
      const Far c = q.top();
      q.pop();
      bool good = true;
      for (const Far& k : keep)
        if (dist(k.id, c.id) < c.d) {
          good = false;
          break;
        }
      if (good) keep.push_back(c);
    }
    for (const Far& k : keep) rs.push({k.d, k.id});
  }

  // Heuristic 1: as heuristic 2, then refill up to NN with the nearest of
  // the rejected candidates.
  void select_h1(NearHeap& rs, int nn) const {
    if (rs.size() < (size_t)nn) return;
    FarHeap q, rejected;
    while (!rs.empty()) {
      q.push({rs.top().d, rs.top().id});
      rs.pop();
    }
    std::vector<Far> keep;
    while (!q.empty()) {
      if (keep.size() >= (size_t)nn) break;
      const Far c = q.top();
      q.pop();
      bool good = true;
      for (const Far& k : keep)
        if (dist(k.id, c.id) < c.d) {
          good = false;
          break;
        }
      if (good) keep.push_back(c);
      else rejected.push(c);
    }
    while (keep.size() < (size_t)nn && !rejected.empty()) {
      keep.push_back(rejected.top());
      rejected.pop();
    }
    for (const Far& k : keep) rs.push({k.d, k.id});
  }

  // Heuristic 3 (hnsw.h:169-258): widen the candidates with their own lists
  // at level L, then take them nearest first into a high tier (no kept
  // neighbour closer) or a second tier (only a passed-over candidate closer),
  // keeping up to NN, high tier first.  The reference collects the widened set
  // in a hash set of node pointers; its iteration order only changes the order
  // among equal distances.  It reads other nodes' lists without their locks,
  // so it runs single-threaded only.
  void select_h3(int32_t self, NearHeap& rs, int nn, int L) {
    std::unordered_set<int32_t> cands;
    while (!rs.empty()) {
      const int32_t t = rs.top().id;
      cands.insert(t);
      for (int32_t f : nodes_[t].links[L]) cands.insert(f);
      rs.pop();
    }
    for (int32_t c : cands)
      if (c != self) rs.push({dist(c, self), c});
    if (rs.size() < (size_t)nn) return;
    std::vector<Near> in(rs.size());
    for (int i = (int)rs.size() - 1; i >= 0; --i) {
      in[i] = rs.top();
      rs.pop();
    }
    std::vector<Near> passed, second, high;
    for (const Near& c : in) {
      if (high.size() >= (size_t)nn) break;
      int good = 2;
      for (const Near& t : passed)
        if (dist(t.id, c.id) < c.d) {
          good = 1;
          break;
        }
      for (const Near& t : high)
        if (dist(t.id, c.id) < c.d) {
          good = 0;
          break;
        }
      if (good)
        for (const Near& t : second)
          if (dist(t.id, c.id) < c.d) {
            good = 0;
            break;
          }
      if (good == 2) high.push_back(c);
      else if (good == 1) second.push_back(c);
      else passed.push_back(c);
    }
    for (const Near& t : high) {
      if (rs.size() >= (size_t)nn) break;
      rs.push(t);
    }
    for (const Near& t : second) {
      if (rs.size() >= (size_t)nn) break;
      rs.push(t);
    }
  }

  // addFriendlevel: append `e` to x's list at level L; when the list is over
  // capacity, re-select all but one with the build's heuristic (the list is
  // then stored farthest first), or drop the farthest for delaunay_type 0.
  void add_friend(int32_t x, int32_t e, int L) {
    Node& nd = nodes_[x];
    std::lock_guard<std::mutex> g(nd.mu);
    auto& lst = nd.links[L];
    if (std::find(lst.begin(), lst.end(), e) != lst.end()) return;
    lst.push_back(e);
    const size_t cap = L > 0 ? p_.maxM : p_.maxM0;
    if (lst.size() <= cap) return;
    if (p_.delaunay > 0) {
      NearHeap rs;
      for (int32_t v : lst) rs.push({dist(x, v), v});
      const int nn = (int)rs.size() - 1;
      if (p_.delaunay == 1) select_h1(rs, nn);
      else if (p_.delaunay == 3) select_h3(x, rs, nn, L);
      else select_h2(rs, nn);
      lst.clear();
      while (!rs.empty()) {
        lst.push_back(rs.top().id);
        rs.pop();
      }
    } else {
      float mx = dist(x, lst[0]);
      size_t at = 0;
      for (size_t i = 1; i < lst.size(); ++i) {
        const float d = dist(x, lst[i]);
        if (d > mx) {
          mx = d;
          at = i;
        }
      }
      lst.erase(lst.begin() + at);
    }
  }

  const int space_;
  const std::vector<const float*>& vec_;
  const size_t dim_;
  const BuildParams p_;
  const size_t n_;
  std::unique_ptr<Node[]> nodes_;
  std::atomic<uint64_t> top_{0};  // maxlevel << 32 | entry point
  std::mutex grow_mu_;            // held by an insertion that may raise maxlevel
};

}  // namespace

float space_distance(int space, const float* a, const float* b, size_t dim) {
  const std::vector<const float*> vec{a, b};
  BuildParams p;
  return Graph(space, vec, dim, p).dist(0, 1);
}

void build_hnsw(int space, const std::vector<std::vector<char>>& objs, const BuildParams& p_in,
                BuiltIndex& out) {
  if (space != kBuildL2 && space != kBuildCosine && space != kBuildNegDot && space != kBuildL1 &&
      space != kBuildLInf)
    throw std::invalid_argument("unknown space code " + std::to_string(space));
  if (p_in.delaunay < 0 || p_in.delaunay > 3)
    throw std::invalid_argument("delaunay_type should be 0, 1, 2 or 3");
  if (p_in.delaunay == 3 && p_in.threads > 1)
    throw std::invalid_argument("delaunay_type 3 needs indexThreadQty=1 (its selection reads other "
                                "nodes' lists while they may change)");
  if (p_in.post < 0 || p_in.post > 2) throw std::invalid_argument("post should be 0, 1 or 2");
  if (p_in.M == 0 || p_in.maxM == 0 || p_in.maxM0 == 0 || p_in.efConstruction == 0)
    throw std::invalid_argument("M, maxM, maxM0 and efConstruction must be positive");
  if (!std::isfinite(p_in.mult) || p_in.mult < 0)
    throw std::invalid_argument("mult must be finite and non-negative (M = 1 needs an explicit mult)");
  const size_t n = objs.size();
  if (n > (size_t)std::numeric_limits<int32_t>::max())
    throw std::invalid_argument("too many data points for 32-bit ids");
  out = BuiltIndex{};
  if (n == 0) return;
  uint64_t len = 0;
  std::vector<const float*> vec(n);
  for (size_t i = 0; i < n; ++i) {
    const auto& o = objs[i];
    if (o.size() < 16) throw std::runtime_error("corrupt data point buffer");
    uint64_t dl;
    std::memcpy(&dl, o.data() + 8, 8);
    if (dl + 16 != o.size() || dl == 0 || dl % 4 != 0)
      throw std::runtime_error("data point " + std::to_string(i) + " is not a dense float vector");
    if (i == 0) len = dl;
    else if (dl != len)
      throw std::runtime_error("data points differ in dimension (" + std::to_string(len / 4) +
                               " and " + std::to_string(dl / 4) + ")");
    vec[i] = reinterpret_cast<const float*>(o.data() + 16);
  }
  const size_t dim = len / 4;
  BuildParams p = p_in;

  Graph g(space, vec, dim, p);
  g.run_pass(false);
  if (p.post == 1 || p.post == 2) {
    std::unique_ptr<Node[]> first = g.take_nodes();
    g.run_pass(true);
    const size_t mx = g.merge_with(first);
    if (p.post == 1 && n > 1) p.maxM0 = mx;
  }

  // optimized layout (hnsw.cc:362-470): [Object buffer | count | ids[maxM0]]
  // per point, unused slots 0x01 bytes; upper levels [count | ids[maxM]]
  const uint64_t data_bytes = 16 + len;
  out.n = (uint32_t)n;
  out.off_l0 = data_bytes;
  out.elem = data_bytes + 4 * (p.maxM0 + 1);
  out.maxM = p.maxM;
  out.maxM0 = p.maxM0;
  out.maxlevel = g.maxlevel();
  out.enter = (uint32_t)g.enter();
  switch (space) {  // optimized: hnsw.cc:381-405; regular: the space's own distance (7, 8, 9)
    case kBuildL2: out.func = p.optimized ? (dim % 16 == 0 ? 1 : 2) : 7; break;
    case kBuildCosine: out.func = p.optimized ? 3 : 8; break;
    case kBuildNegDot: out.func = p.optimized ? 4 : 9; break;
    case kBuildL1: out.func = 5; break;
    default: out.func = 6; break;
  }
  out.block.assign(n * out.elem, 1);
  out.sizemass.assign(n, 0);
  out.upper_off.assign(n, -1);
  for (size_t i = 0; i < n; ++i) {
    char* e = out.block.data() + i * out.elem;
    std::memcpy(e, objs[i].data(), data_bytes);
    if (space == kBuildCosine && p.optimized) normalize_vect(reinterpret_cast<float*>(e + 16), dim);
    const Node& nd = g.node(i);
    const auto& l0 = nd.links[0];
    if (l0.size() > p.maxM0) throw std::runtime_error("internal error: level-0 list over capacity");
    const int32_t cnt = (int32_t)l0.size();
    std::memcpy(e + data_bytes, &cnt, 4);
    if (cnt) std::memcpy(e + data_bytes + 4, l0.data(), 4 * l0.size());
    if (nd.level >= 1) {
      out.sizemass[i] = (uint32_t)(nd.level * (p.maxM + 1) * 4);
      out.upper_off[i] = (int32_t)out.upper.size();
      for (int L = 1; L <= nd.level; ++L) {
        const auto& lst = nd.links[L];
        if (lst.size() > p.maxM) throw std::runtime_error("internal error: upper list over capacity");
        out.upper.push_back((int32_t)lst.size());
        out.upper.insert(out.upper.end(), lst.begin(), lst.end());
        out.upper.resize(out.upper.size() + (p.maxM - lst.size()), 0);
      }
    }
  }
}

}  // namespace nmsamd

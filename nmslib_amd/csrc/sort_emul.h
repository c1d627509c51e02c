// Order-exact sorting helpers shared by host and device code.
//
// SearchV1Merge sorts its per-expansion buffer with std::sort
// (similarity_search/src/method/hnsw_distfunc_opt.cc:251) on Items compared by
// key only (include/sort_arr_bi.h:42-44), and SearchOld keeps its queues in
// std::priority_queue (hnsw_distfunc_opt.cc:99-102).  When keys are equal the
// order those produce is a property of the libstdc++ algorithms, so the
// deterministic search mode replays the same algorithm steps:
//   * introsort: recursion on the right part while the range holds more than
//     16 items, pivot = median of (second, middle, last) swapped to the front,
//     unguarded Hoare partition; heap sort once the depth budget
//     2*floor(log2 n) is spent; a final insertion sort (guarded on the first
//     16 items, unguarded after them);
//   * binary max-heaps with the usual sift-down-to-leaf-then-sift-up pop.
// Only the comparison "key <" is used, as in the reference.
#ifndef NMSLIB_AMD_SORT_EMUL_H
#define NMSLIB_AMD_SORT_EMUL_H

#if defined(__HIPCC__)
#define NMS_HD __host__ __device__
#else
#define NMS_HD
#endif

namespace nmsamd {

template <class K, class V>
struct KVRef {
  K* k;
  V* v;
  NMS_HD void swap(int a, int b) const {
    K tk = k[a];
    k[a] = k[b];
    k[b] = tk;
    V tv = v[a];
    v[a] = v[b];
    v[b] = tv;
  }
  NMS_HD bool less(int a, int b) const { return k[a] < k[b]; }
};

template <class K, class V>
NMS_HD inline void emu_heap_sift(KVRef<K, V> r, int base, int hole, int len, K vk, V vv) {
  // sift the hole down to a leaf along the larger child, then back up
  const int top = hole;
  int child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (r.k[base + child] < r.k[base + child - 1]) --child;
    r.k[base + hole] = r.k[base + child];
    r.v[base + hole] = r.v[base + child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    r.k[base + hole] = r.k[base + child - 1];
    r.v[base + hole] = r.v[base + child - 1];
    hole = child - 1;
  }
  int parent = (hole - 1) / 2;
  while (hole > top && r.k[base + parent] < vk) {
    r.k[base + hole] = r.k[base + parent];
    r.v[base + hole] = r.v[base + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  r.k[base + hole] = vk;
  r.v[base + hole] = vv;
}

template <class K, class V>
NMS_HD inline void emu_heap_sort(KVRef<K, V> r, int lo, int hi) {
  const int len = hi - lo;
  if (len >= 2) {
    for (int parent = (len - 2) / 2;; --parent) {
      emu_heap_sift(r, lo, parent, len, r.k[lo + parent], r.v[lo + parent]);
      if (parent == 0) break;
    }
  }
  for (int last = hi; last - lo > 1;) {
    --last;
    K vk = r.k[last];
    V vv = r.v[last];
    r.k[last] = r.k[lo];
    r.v[last] = r.v[lo];
    emu_heap_sift(r, lo, 0, last - lo, vk, vv);
  }
}

template <class K, class V>
NMS_HD inline void emu_linear_insert(KVRef<K, V> r, int pos) {
  K vk = r.k[pos];
  V vv = r.v[pos];
  int next = pos - 1;
  while (vk < r.k[next]) {
    r.k[pos] = r.k[next];
    r.v[pos] = r.v[next];
    pos = next;
    --next;
  }
  r.k[pos] = vk;
  r.v[pos] = vv;
}

template <class K, class V>
NMS_HD inline void emu_insertion_sort(KVRef<K, V> r, int lo, int hi) {
  if (lo == hi) return;
  for (int i = lo + 1; i < hi; ++i) {
    if (r.k[i] < r.k[lo]) {
      K vk = r.k[i];
      V vv = r.v[i];
      for (int j = i; j > lo; --j) {
        r.k[j] = r.k[j - 1];
        r.v[j] = r.v[j - 1];
      }
      r.k[lo] = vk;
      r.v[lo] = vv;
    } else {
      emu_linear_insert(r, i);
    }
  }
}

template <class K, class V>
NMS_HD inline int emu_partition_pivot(KVRef<K, V> r, int lo, int hi) {
  const int mid = lo + (hi - lo) / 2;
  const int a = lo + 1, b = mid, c = hi - 1;
  if (r.less(a, b)) {
    if (r.less(b, c)) r.swap(lo, b);
    else if (r.less(a, c)) r.swap(lo, c);
    else r.swap(lo, a);
  } else if (r.less(a, c)) {
    r.swap(lo, a);
  } else if (r.less(b, c)) {
    r.swap(lo, c);
  } else {
    r.swap(lo, b);
  }
  int first = lo + 1, last = hi;
  for (;;) {
    while (r.less(first, lo)) ++first;
    --last;
    while (r.less(lo, last)) --last;
    if (!(first < last)) return first;
    r.swap(first, last);
    ++first;
  }
}

// Sorts [0, n) by key exactly as std::sort orders it.  The introsort's right
// parts are independent of its left parts, so an explicit stack replaces the
// recursion; each part keeps the depth budget it would have had.  The stack
// holds at most 2*depth+1 frames (n <= 4096 here).
template <class K, class V>
NMS_HD inline void emu_std_sort(K* key, V* val, int n) {
  if (n <= 1) return;
  KVRef<K, V> r{key, val};
  int lg = 0;
  while ((2 << lg) <= n) ++lg;
  struct Frame { int lo, hi, depth; };
  Frame stack[56];
  int sp = 0;
  stack[sp++] = Frame{0, n, 2 * lg};
  while (sp > 0) {
    const Frame f = stack[--sp];
    if (f.hi - f.lo <= 16) continue;
    if (f.depth == 0) {
      emu_heap_sort(r, f.lo, f.hi);
      continue;
    }
    const int cut = emu_partition_pivot(r, f.lo, f.hi);
    stack[sp++] = Frame{f.lo, cut, f.depth - 1};
    stack[sp++] = Frame{cut, f.hi, f.depth - 1};
  }
  if (n > 16) {
    emu_insertion_sort(r, 0, 16);
    for (int i = 16; i < n; ++i) emu_linear_insert(r, i);
  } else {
    emu_insertion_sort(r, 0, n);
  }
}

// std::priority_queue<T> (max-heap on operator<) push / pop over arrays.
template <class K, class V>
NMS_HD inline void emu_heap_push(K* key, V* val, int len /* after append */) {
  int hole = len - 1;
  K vk = key[hole];
  V vv = val[hole];
  int parent = (hole - 1) / 2;
  while (hole > 0 && key[parent] < vk) {
    key[hole] = key[parent];
    val[hole] = val[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  key[hole] = vk;
  val[hole] = vv;
}

template <class K, class V>
NMS_HD inline void emu_heap_pop(K* key, V* val, int len /* before pop */) {
  if (len <= 1) return;
  KVRef<K, V> r{key, val};
  const int last = len - 1;
  K vk = key[last];
  V vv = val[last];
  key[last] = key[0];
  val[last] = val[0];
  emu_heap_sift(r, 0, 0, last, vk, vv);
}

}  // namespace nmsamd

#endif

// Test: nmslib_amd/csrc/sort_emul.h reproduces the element order of
// libstdc++ std::sort and std::priority_queue when keys tie.
#include "../../nmslib_amd/csrc/sort_emul.h"
#include <algorithm>
#include <cstdio>
#include <queue>
#include <random>
#include <vector>
struct It { float key; int data; bool operator<(const It& o) const { return key < o.key; } };
int main() {
  std::mt19937 rng(1);
  long bad = 0, tot = 0;
  for (int trial = 0; trial < 200000; ++trial) {
    int n = 1 + rng() % (trial % 7 == 0 ? 700 : 70);
    int vals = 1 + rng() % 12;
    std::vector<It> a(n);
    for (int i = 0; i < n; ++i) a[i] = {float(rng() % vals), i};
    std::vector<float> k(n); std::vector<int> v(n);
    for (int i = 0; i < n; ++i) { k[i] = a[i].key; v[i] = a[i].data; }
    std::sort(a.begin(), a.end());
    nmsamd::emu_std_sort(k.data(), v.data(), n);
    for (int i = 0; i < n; ++i) if (a[i].data != v[i] || a[i].key != k[i]) { bad++; break; }
    tot++;
  }
  printf("sort: %ld / %ld mismatched\n", bad, tot);
  // priority queue
  long pbad = 0;
  for (int trial = 0; trial < 20000; ++trial) {
    std::priority_queue<It> pq; std::vector<float> k; std::vector<int> v;
    int ops = 1 + rng() % 300;
    for (int o = 0; o < ops; ++o) {
      if (!pq.empty() && rng() % 3 == 0) {
        It t = pq.top(); pq.pop();
        if (t.data != v[0] || t.key != k[0]) { pbad++; break; }
        nmsamd::emu_heap_pop(k.data(), v.data(), (int)k.size()); k.pop_back(); v.pop_back();
      } else {
        It t{float(rng() % 5), o}; pq.push(t); k.push_back(t.key); v.push_back(t.data);
        nmsamd::emu_heap_push(k.data(), v.data(), (int)k.size());
      }
    }
  }
  printf("heap: %ld mismatched\n", pbad);
  return bad || pbad;
}

// CPU check of the register-resident exact insertion algorithm used by the
// kernel's tie path, against the plain sequential SortArrBI semantics.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
struct It { float k; int id; };
// reference semantics
size_t ref_insert(std::vector<It>& v, size_t& n, size_t C, float key, int id) {
  size_t cur = n - 1;
  if (v[cur].k <= key) { if (n < C) { v[n] = {key, id}; return n++; } return n; }
  size_t prev = cur, step = 1;
  while (cur > 0 && v[cur].k > key) { prev = cur; cur -= step; step *= 2; if (step > cur) step = cur; }
  if (cur < prev) { size_t lo = cur, hi = prev; while (lo < hi) { size_t mid = (lo + hi) / 2; if (v[mid].k < key) lo = mid + 1; else hi = mid; } cur = lo; }
  if (n < C) ++n;
  for (size_t i = n - 1; i > cur; --i) v[i] = v[i - 1];
  v[cur] = {key, id};
  return cur;
}
int main() {
  std::mt19937 rng(7);
  long bad = 0;
  for (int trial = 0; trial < 300000; ++trial) {
    int C = 1 + rng() % 40, n0 = 1 + rng() % C, m = 1 + rng() % 40, vals = 1 + rng() % 15;
    std::vector<It> A(C + 1);
    std::vector<float> ak(n0);
    for (auto& x : ak) x = float(rng() % vals);
    std::sort(ak.begin(), ak.end());
    for (int i = 0; i < n0; ++i) A[i] = {ak[i], 1000 + i};
    std::vector<It> B(m);
    std::vector<float> bk(m);
    for (auto& x : bk) x = float(rng() % vals);
    std::sort(bk.begin(), bk.end());
    for (int j = 0; j < m; ++j) B[j] = {bk[j], 5000 + j};
    // reference
    std::vector<It> R = A; size_t rn = n0; size_t rce = rng() % (n0 + 1), ce0 = rce;
    for (int j = 0; j < m; ++j) { size_t r = ref_insert(R, rn, C, B[j].k, B[j].id); if (r < rce) rce = r; }
    // virtual algorithm
    auto lower = [&](float b) { int c = 0; for (int i = 0; i < n0; ++i) c += A[i].k < b; return c; };
    auto upper = [&](float b) { int c = 0; for (int i = 0; i < n0; ++i) c += A[i].k <= b; return c; };
    std::vector<float> sk; std::vector<int> sid, sab;
    int na = n0; size_t ce = ce0;
    for (int j = 0; j < m; ++j) {
      float b = B[j].k; int LA = std::min(lower(b), na), UA = std::min(upper(b), na);
      int ns = sk.size(), nv = na + ns;
      bool last_s = ns > 0 && sab[ns - 1] == na;
      float lastkey = last_s ? sk[ns - 1] : A[na - 1].k;
      int ret;
      if (lastkey <= b) {
        if (nv < C) { sk.push_back(b); sab.push_back(na); sid.push_back(B[j].id); }
        ret = nv;
      } else {
        int LS = 0, US = 0; for (int t = 0; t < ns; ++t) { LS += sk[t] < b; US += sk[t] <= b; }
        int L = LA + LS, U = UA + US;
        int c = nv - 1, prev = c, step = 1;
        while (c > 0 && c >= U) { prev = c; c -= step; step *= 2; if (step > c) step = c; }
        int pos = c < prev ? std::max(c, L) : c;
        if (nv == C) { if (last_s) { sk.pop_back(); sab.pop_back(); sid.pop_back(); } else --na; }
        ns = sk.size();
        int t0 = 0; for (int t = 0; t < ns; ++t) t0 += (sab[t] + t < pos);
        sk.insert(sk.begin() + t0, b); sab.insert(sab.begin() + t0, pos - t0); sid.insert(sid.begin() + t0, B[j].id);
        ret = pos;
      }
      if ((size_t)ret < ce) ce = ret;
    }
    std::vector<It> V(C + 1); int ns = sk.size();
    for (int i = 0; i < na; ++i) { int before = 0; for (int t = 0; t < ns; ++t) before += sab[t] <= i; V[i + before] = A[i]; }
    for (int t = 0; t < ns; ++t) V[sab[t] + t] = {sk[t], sid[t]};
    size_t vn = na + ns;
    bool ok = vn == rn && ce == rce;
    for (size_t i = 0; ok && i < rn; ++i) ok = V[i].k == R[i].k && V[i].id == R[i].id;
    if (!ok) { if (bad < 3) printf("mismatch trial %d C=%d n0=%d m=%d\n", trial, C, n0, m); bad++; }
  }
  printf("virtual insertion: %ld mismatches\n", bad);
  return bad != 0;
}

// Exact k-NN by brute force, second stage: the matrix-core screen
// (bruteforce.hip) keeps the kc best candidates of every query by
// |x|^2 + |q|^2 - 2 q.x (or 1 - q.x, -q.x); here every candidate's distance is
// recomputed in FP32 in the reference's own summation order and the
// candidates are ranked by (distance, internal id), which is how the
// reference's seq_search orders its answers (seqsearch.cc:143-177 through
// KNNQueue, knnqueue.h:55-81; gold_standard.h:73-76).
//
// A certificate decides whether the screen could have dropped an answer: with
// e the bound on the screen's error (or 0 where the screen is exact, integer
// data), every point outside the kc candidates has a screen distance >= Dmax
// (the kc-th), hence an exact distance >= Dmax - e.  If Dmax - 2e exceeds the
// k-th exact distance (mapped to the screen's scale), none of them can enter
// the top k, not even on a tie.  Queries that fail it are listed and searched
// again by an exact scan over all points (bf_exact_scan_kernel).
#include "hnsw_search_impl.h"

namespace nmsamd {

struct RescoreArgs {
  DevIndex ix;
  const float* queries;   // [nq][dstride], as the screen saw them (normalised for cosine)
  int64_t nq;
  int kc;                 // screened candidates per query (<= 256)
  const float* cand_d;    // [nq][kc] screen distances, ascending
  const int32_t* cand_i;  // [nq][kc] internal ids (>= n: none)
  const float* eps;       // [nq] bound on the screen's error (0: exact screen)
  int k;
  int sqrt_out;           // the exact distance is the square root of the screen's scale
  int32_t* ids;           // [nq][k] object ids
  float* dists;           // [nq][k]
  int32_t* counts;        // [nq]
  int32_t* flagged;       // [nq] queries that need the exact scan
  int* nflag;
};

template <int D, int FUNC>
__global__ void __launch_bounds__(64) bf_rescore_kernel(RescoreArgs a) {
  __shared__ __align__(16) float qs[2 * 512];
  __shared__ float ed[256];
  __shared__ uint32_t ei[256];
  const int lane = threadIdx.x;
  const int h = lane & 1;
  const int64_t qi = blockIdx.x;
  if (qi >= a.nq) return;
  const DevIndex& ix = a.ix;
  for (int c = lane; c < ix.dstride; c += 64) qs[c] = a.queries[qi * ix.dstride + c];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  const float* cd = a.cand_d + qi * a.kc;
  const int32_t* ci = a.cand_i + qi * a.kc;
  int nvalid = 0;
  for (int b = 0; b < a.kc; b += 64)
    nvalid += __popcll(__ballot(b + lane < a.kc && (uint32_t)ci[b + lane] < (uint32_t)ix.n));
  // exact distances, a lane pair per candidate (the search's distance code)
  for (int b = 0; b < 256; b += 32) {
    const int s = b + (lane >> 1);
    const bool ok = s < nvalid;
    const int id = ok ? ci[s] : 0;
    const float dd = (b < nvalid) ? dist_f32<D, FUNC>(qs, ix.vec + (int64_t)id * ix.dstride, h, ix.dim)
                                  : 0.f;
    if (h == 0) {
      ed[s] = ok ? dd : kInf();
      ei[s] = ok ? (uint32_t)id : 0xffffffffu;
    }
  }
  __syncthreads();
  float key[4];
  uint32_t id[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    key[r] = ed[r * 64 + lane];
    id[r] = ei[r * 64 + lane];
  }
  bitonic_sort<4>(key, id, lane);
  const int kk = min(a.k, nvalid);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = r * 64 + lane;
    if (j < kk) {
      a.ids[qi * a.k + j] = ix.ext_id[id[r]];
      a.dists[qi * a.k + j] = key[r];
    } else if (j < a.k) {
      a.ids[qi * a.k + j] = -1;
      a.dists[qi * a.k + j] = 0.f;
    }
  }
  // the certificate: every point was a candidate, or the kc-th screen
  // distance clears the k-th exact one by twice the screen's error
  bool sure = nvalid < a.kc || (int64_t)a.kc >= ix.n;
  if (!sure && kk == a.k) {
    const int j = a.k - 1;
    float t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(key[j >> 6]), j & 63));
    // a square root of FP32 keeps equal outputs for inputs up to 3 ulp apart
    if (a.sqrt_out) t = t * t * (1.0f + 4.0f * 5.9604645e-8f);
    const float e = a.eps[qi];
    sure = cd[a.kc - 1] - 2.0f * e > t;
  }
  if (lane == 0) {
    a.counts[qi] = sure ? kk : 0;
    if (!sure) a.flagged[atomicAdd(a.nflag, 1)] = (int32_t)qi;
  }
}

// The exact scan of listed queries: block (chunk, listed query) walks its
// chunk of the base vectors, a lane pair per vector, and keeps the chunk's k
// best by (distance, internal id) in LDS; lane 0 inserts the few distances
// that beat the current k-th.  The chunk lists are merged by bf_merge_kernel.
struct ScanArgs {
  DevIndex ix;
  const float* queries;
  const int32_t* rows;    // [nrows] query of each listed row
  const int* nrows;
  int k, nchunks;
  int64_t chunk_len;
  float* part_d;          // [nrows_max][nchunks][k]
  int32_t* part_i;
};

template <int D, int FUNC>
__global__ void __launch_bounds__(64) bf_exact_scan_kernel(ScanArgs a) {
  __shared__ __align__(16) float qs[2 * 512];
  __shared__ float ld[128];
  __shared__ int li[128];
  __shared__ float nd[32];
  __shared__ int ni[32];
  const int lane = threadIdx.x;
  const int h = lane & 1;
  const int row = blockIdx.y;
  if (row >= *a.nrows) return;
  const int chunk = blockIdx.x;
  const DevIndex& ix = a.ix;
  const int64_t qi = a.rows[row];
  for (int c = lane; c < ix.dstride; c += 64) qs[c] = a.queries[qi * ix.dstride + c];
  __syncthreads();
  const int64_t lo = (int64_t)chunk * a.chunk_len;
  const int64_t hi = min(ix.n, lo + a.chunk_len);
  int cnt = 0;
  float thr = kInf();
  int thr_i = 0x7fffffff;
  for (int64_t b = lo; b < hi; b += 32) {
    const int64_t x = b + (lane >> 1);
    const bool ok = x < hi;
    const float dd = dist_f32<D, FUNC>(qs, ix.vec + (ok ? x : lo) * ix.dstride, h, ix.dim);
    const bool take = ok && h == 0 &&
                      (cnt < a.k || dd < thr || (dd == thr && (int)x < thr_i));
    const uint64_t m = __ballot(take);
    if (m == 0) continue;
    if (take) {
      const int p = __popcll(m & ((1ull << lane) - 1));
      nd[p] = dd;
      ni[p] = (int)x;
    }
    __syncthreads();
    if (lane == 0) {
      const int nn = __popcll(m);
      for (int j = 0; j < nn; ++j) {
        const float d = nd[j];
        const int id = ni[j];
        if (cnt == a.k && !(d < ld[cnt - 1] || (d == ld[cnt - 1] && id < li[cnt - 1]))) continue;
        int p = cnt < a.k ? cnt++ : a.k - 1;
        while (p > 0 && (d < ld[p - 1] || (d == ld[p - 1] && id < li[p - 1]))) {
          ld[p] = ld[p - 1];
          li[p] = li[p - 1];
          --p;
        }
        ld[p] = d;
        li[p] = id;
      }
    }
    __syncthreads();
    cnt = __shfl(cnt, 0);
    thr = cnt == a.k ? ld[a.k - 1] : kInf();
    thr_i = cnt == a.k ? li[a.k - 1] : 0x7fffffff;
  }
  cnt = __shfl(cnt, 0);
  const int64_t off = ((int64_t)row * a.nchunks + chunk) * a.k;
  for (int j = lane; j < a.k; j += 64) {
    a.part_d[off + j] = j < cnt ? ld[j] : kInf();
    a.part_i[off + j] = j < cnt ? li[j] : 0x7fffffff;
  }
}

using RescoreFn = void (*)(RescoreArgs);
using ScanFn = void (*)(ScanArgs);

template <int D>
static void pick(int func, RescoreFn* r, ScanFn* s) {
  switch (func) {
#define NMS_F(F)                          \
  case F:                                 \
    *r = bf_rescore_kernel<D, F>;         \
    *s = bf_exact_scan_kernel<D, F>;      \
    break;
    NMS_F(kL2Sqr16) NMS_F(kL2Sqr) NMS_F(kNormCos) NMS_F(kNegDot)
    NMS_F(kSpaceL2) NMS_F(kSpaceCos) NMS_F(kSpaceNegDot)
#undef NMS_F
    default:
      *r = nullptr;
      *s = nullptr;
  }
}

static void pick_fns(int dim, int func, RescoreFn* r, ScanFn* s) {
  // the AVX-order kernels are specialised by dimension, the SSE-order ones
  // take it at run time
  if (func >= kSpaceL2) return pick<0>(func, r, s);
  switch (dim) {
    case 96: return pick<96>(func, r, s);
    case 100: return pick<100>(func, r, s);
    case 128: return pick<128>(func, r, s);
    default: return pick<0>(func, r, s);
  }
}

hipError_t launch_brute_rescore(const DevIndex& ix, const float* queries, int64_t nq, int kc,
                                const float* cand_d, const int32_t* cand_i, const float* eps,
                                int k, int func, int sqrt_out, int32_t* ids, float* dists,
                                int32_t* counts, int32_t* flagged, int* nflag,
                                hipStream_t stream) {
  RescoreFn r;
  ScanFn s;
  pick_fns(ix.dim, func, &r, &s);
  if (!r || kc > 256 || k > kc || ix.dstride > 512) return hipErrorInvalidValue;
  RescoreArgs a{ix, queries, nq, kc, cand_d, cand_i, eps, k, sqrt_out, ids, dists, counts,
                flagged, nflag};
  hipLaunchKernelGGL(r, dim3((unsigned)nq), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_brute_scan(const DevIndex& ix, const float* queries, const int32_t* rows,
                             const int* nrows, int nrows_max, int k, int func, int nchunks,
                             float* part_d, int32_t* part_i, hipStream_t stream) {
  RescoreFn r;
  ScanFn s;
  pick_fns(ix.dim, func, &r, &s);
  if (!s || k > 128 || ix.dstride > 512 || nrows_max <= 0) return hipErrorInvalidValue;
  ScanArgs a{ix, queries, rows, nrows, k, nchunks, (ix.n + nchunks - 1) / nchunks, part_d, part_i};
  hipLaunchKernelGGL(s, dim3((unsigned)nchunks, (unsigned)nrows_max), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace nmsamd

// Exact k-NN by brute force, first stage (the screen) on MFMA: the device
// counterpart of the reference's seq_search method
// (similarity_search/src/method/seqsearch.cc:143-177) is this screen followed
// by the exact re-scoring of bruteforce_exact.hip.  It is the ground truth for
// workloads too large for the host and the dense-contraction kernel of the
// path.
//
// Kernel 1 (bf_tile_kernel): a wave owns 32 queries and one chunk of the base
// vectors.  Per tile of 32 base vectors it computes the 32x32 block of dot
// products with v_mfma_f32_32x32x2_f32 -- K slot 0 carries dimensions 0..63,
// slot 1 dimensions 64..127, so each lane streams one contiguous half-row of
// its base vector and keeps its query half-row in registers -- turns them into
// distances (l2: |x|^2 + |q|^2 - 2 q.x, cosine: 1 - q.x on unit vectors,
// negdotprod: -q.x) and keeps a per-lane top-k in LDS.  Each query's two lanes
// write their lists for the chunk.
// Kernel 2 (bf_merge_kernel): one wave per query merges the sorted partial
// lists by (distance, index) into the query's kc candidates.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hnsw_device.h"

namespace nmsamd {

typedef float floatx16 __attribute__((ext_vector_type(16)));

struct BruteArgs {
  const float* vec;       // [n][dstride]
  const float* xnorm;     // [n] |x|^2 (l2)
  const float* queries;   // [nq][dstride] (normalised for cosine)
  int64_t n, nq;
  int dstride, func, k, nchunks;
  int64_t chunk_len;
  float* part_d;          // [nq][nchunks*2][k]
  int32_t* part_i;
};

__device__ __forceinline__ bool pair_less(float da, int ia, float db, int ib) {
  return da < db || (da == db && ia < ib);
}

__global__ void __launch_bounds__(64) bf_tile_kernel(BruteArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int ncol = lane & 31;       // the query this lane's accumulators belong to
  const int half = lane >> 5;       // K slot: dimensions [64*half, 64*half + 64)
  const int64_t q0 = (int64_t)blockIdx.x * 32;
  const int chunk = blockIdx.y;
  const int K = a.k;
  float* kd = reinterpret_cast<float*>(smem);          // [64][K] per-lane top-k
  int* ki = reinterpret_cast<int*>(kd + 64 * K);
  float* xn = reinterpret_cast<float*>(ki + 64 * K);   // [32] norms of the tile

  // query half-row in registers (zero past dstride / past nq)
  float qr[64];
  const int64_t qrow = q0 + ncol;
  float qq = 0.f;
#pragma unroll
  for (int t = 0; t < 64; t += 4) {
    const int dcol = 64 * half + t;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (qrow < a.nq && dcol < a.dstride)
      v = *reinterpret_cast<const float4*>(a.queries + qrow * a.dstride + dcol);
    qr[t] = v.x; qr[t + 1] = v.y; qr[t + 2] = v.z; qr[t + 3] = v.w;
    qq += v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w;
  }
  qq += __shfl_xor(qq, 32);  // both halves of the row

  int cnt = 0;
  float thr = __builtin_inff();
  int thr_i = 0x7fffffff;
  const int64_t lo = (int64_t)chunk * a.chunk_len;
  const int64_t hi = min(a.n, lo + a.chunk_len);
  for (int64_t base = lo; base < hi; base += 32) {
    const int64_t xrow = base + ncol;
    const bool xin = xrow < hi;
    float4 xr[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int dcol = 64 * half + 4 * t;
      xr[t] = (xin && dcol < a.dstride)
                  ? *reinterpret_cast<const float4*>(a.vec + xrow * a.dstride + dcol)
                  : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (half == 0) xn[ncol] = (a.func == kL2Sqr16 || a.func == kL2Sqr) && xin ? a.xnorm[xrow] : 0.f;
    floatx16 acc = {};
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[t].x, qr[4 * t], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[t].y, qr[4 * t + 1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[t].z, qr[4 * t + 2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[t].w, qr[4 * t + 3], acc, 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // accumulator j of this lane is row (j/4)*8 + half*4 + j%4 of the tile
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int row = (j >> 2) * 8 + half * 4 + (j & 3);
      const int64_t xi = base + row;
      if (xi >= hi || qrow >= a.nq) continue;
      const float dot = acc[j];
      float d;
      if (a.func == kL2Sqr16 || a.func == kL2Sqr) {
        d = xn[row] + qq - 2.0f * dot;
        d = d < 0.f ? 0.f : d;
      } else if (a.func == kNormCos) {
        const float c = dot < -1.f ? -1.f : (dot > 1.f ? 1.f : dot);
        d = 1.0f - c;
        d = d < 0.f ? 0.f : d;
      } else {
        d = -dot;
      }
      const int id = (int)xi;
      if (cnt < K || pair_less(d, id, thr, thr_i)) {
        // insert into this lane's sorted list
        int p = cnt < K ? cnt : K - 1;
        if (cnt < K) ++cnt;
        float* ld = kd + lane * K;
        int* li = ki + lane * K;
        while (p > 0 && pair_less(d, id, ld[p - 1], li[p - 1])) {
          ld[p] = ld[p - 1];
          li[p] = li[p - 1];
          --p;
        }
        ld[p] = d;
        li[p] = id;
        if (cnt == K) {
          thr = ld[K - 1];
          thr_i = li[K - 1];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  if (qrow < a.nq) {
    const int64_t off = (qrow * (a.nchunks * 2) + chunk * 2 + half) * K;
    for (int j = 0; j < K; ++j) {
      a.part_d[off + j] = j < cnt ? kd[lane * K + j] : __builtin_inff();
      a.part_i[off + j] = j < cnt ? ki[lane * K + j] : 0x7fffffff;
    }
  }
}

// one wave per list group: k-way merge of nlists sorted lists of length k by
// (distance, internal id) into output row row_of[g] (g when row_of is null);
// ids are object ids through ext_id, or internal ids when ext_id is null
__global__ void __launch_bounds__(64) bf_merge_kernel(const float* part_d, const int32_t* part_i,
                                                      int nlists, int k, int64_t ngroups,
                                                      const int32_t* row_of,
                                                      const int32_t* ext_id, int32_t* ids,
                                                      float* dists, int32_t* counts) {
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x;
  if (g >= ngroups) return;
  const int64_t q = row_of ? row_of[g] : g;
  const float* pd = part_d + g * nlists * k;
  const int32_t* pi = part_i + g * nlists * k;
  int head = 0;  // this lane's position in list `lane`
  int found = 0;
  for (int r = 0; r < k; ++r) {
    float d = __builtin_inff();
    int i = 0x7fffffff;
    if (lane < nlists && head < k) {
      d = pd[lane * k + head];
      i = pi[lane * k + head];
    }
    float bd = d;
    int bi = i, bl = lane;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const float od = __shfl_xor(bd, o);
      const int oi = __shfl_xor(bi, o);
      const int ol = __shfl_xor(bl, o);
      if (pair_less(od, oi, bd, bi)) {
        bd = od;
        bi = oi;
        bl = ol;
      }
    }
    if (bi == 0x7fffffff) break;
    if (lane == bl) ++head;
    if (lane == 0) {
      ids[q * k + r] = ext_id ? ext_id[bi] : bi;
      dists[q * k + r] = bd;
    }
    ++found;
  }
  for (int r = found + lane; r < k; r += 64) {
    ids[q * k + r] = -1;
    dists[q * k + r] = 0.f;
  }
  if (lane == 0 && counts) counts[q] = found;
}

__global__ void bf_norm_kernel(const float* vec, float* out, int64_t n, int dstride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < dstride; ++c) s += vec[i * dstride + c] * vec[i * dstride + c];
  out[i] = s;
}

hipError_t launch_brute_norms(const float* vec, float* out, int64_t n, int dstride,
                              hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(bf_norm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, vec,
                     out, n, dstride);
  return hipGetLastError();
}

// The screen: kc candidates per query by the matrix-core distance, merged
// into cand_d / cand_i ([nq][kc], internal ids, ascending by (distance, id)).
hipError_t launch_brute_screen(const float* vec, const float* xnorm, const float* queries,
                               int64_t n, int64_t nq, int dstride, int func, int kc, int nchunks,
                               float* part_d, int32_t* part_i, float* cand_d, int32_t* cand_i,
                               hipStream_t stream) {
  if (dstride > 128 || kc < 1 || kc > 256 || nchunks * 2 > 64) return hipErrorInvalidValue;
  BruteArgs a{vec, xnorm, queries, n, nq, dstride, func, kc, nchunks,
              (n + nchunks - 1) / nchunks, part_d, part_i};
  a.chunk_len = (a.chunk_len + 31) / 32 * 32;
  const size_t lds = (size_t)64 * kc * 8 + 32 * 4;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(bf_tile_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(bf_tile_kernel, dim3((unsigned)((nq + 31) / 32), (unsigned)nchunks), dim3(64),
                     lds, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bf_merge_kernel, dim3((unsigned)nq), dim3(64), 0, stream, part_d, part_i,
                     nchunks * 2, kc, nq, nullptr, nullptr, cand_i, cand_d, nullptr);
  return hipGetLastError();
}

hipError_t launch_brute_merge(const float* part_d, const int32_t* part_i, int nlists, int k,
                              int64_t rows, const int32_t* row_of, const int32_t* ext_id,
                              int32_t* ids, float* dists, int32_t* counts, hipStream_t stream) {
  if (nlists < 1 || nlists > 64 || rows <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(bf_merge_kernel, dim3((unsigned)rows), dim3(64), 0, stream, part_d, part_i,
                     nlists, k, rows, row_of, ext_id, ids, dists, counts);
  return hipGetLastError();
}

}  // namespace nmsamd

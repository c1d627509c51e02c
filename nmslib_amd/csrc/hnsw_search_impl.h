// HNSW k-NN search kernels for CDNA4 (gfx950): one wavefront per query.
// (Header of device code; search_f32.hip / search_bf16.hip / search_fp8.hip
// instantiate it per vector storage type.)
//
// Semantics follow the reference's optimized search
// (similarity_search/src/method/hnsw_distfunc_opt.cc):
//   * greedy descent over the upper levels (:168-198): a node moves to a
//     strictly closer neighbour, scanning the list in stored order;
//   * SearchV1Merge at level 0 (:200-281): a bounded sorted array of
//     capacity max(ef, k), one candidate expanded per step, new items sorted
//     and inserted with push_or_replace_non_empty_exp (sort_arr_bi.h:159-199);
//   * SearchOld (:46-150) with its two binary heaps, for ef >= 1000 or
//     algoType=old;
//   * the first min(k, size) items reported in (distance, object) order
//     (knnqueue.h:73-74; objects are ordered by their internal index).
//
// MI355X mapping:
//   * a lane pair owns one candidate vector: lane h of the pair loads float4s
//     [16c+4h, 16c+4h+4) and [16c+8+4h, ...) of each 16-float chunk, i.e. the
//     four AVX accumulator lanes 4h..4h+3 of the reference kernel, so the
//     FP32 distance is computed in the reference's exact summation order
//     (hnsw_distfunc_opt_impl_inline.h:42-173) while 32 vectors stream per
//     pass (16 KB of 128-d vectors in flight per wave);
//   * the sorted array, the query and the per-expansion buffers live in LDS;
//     new items are merged in parallel (rank by binary search) and only an
//     expansion whose keys tie an existing key falls back to item-by-item
//     insertion that replays the reference's exponential-search window;
//   * the visited set is the reference's epoch-byte array, one per resident
//     wave, in HBM;
//   * the grid is persistent: waves pull queries from an atomic counter, so
//     queries of uneven ef / k balance across the device.
//
// This file is compiled with -ffp-contract=off: FP32 distances must not be
// contracted into FMAs (the reference rounds the product before the add).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hnsw_device.h"
#include "sort_emul.h"

namespace nmsamd {

#define WSYNC()                                               \
  do {                                                        \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");    \
    __builtin_amdgcn_wave_barrier();                          \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");    \
  } while (0)

__device__ __forceinline__ float kInf() { return __builtin_inff(); }

// A node id read back from the LDS search state just before it addresses
// global memory.  An id outside [0, n) could only come from a corrupted
// state: it is reported through the error word (bit 2, raised as an error by
// the host) and replaced by node 0, so no access leaves the index.
__device__ __forceinline__ int checked_node(int id, int n, int* err) {
  if ((unsigned)id < (unsigned)n) return id;
  if (err) atomicOr(err, 2);
  return 0;
}

// The visited marks of one slot (hnsw.h:568-591: the reference's per-query
// byte array with a version counter).  Bytes: one byte per node holding the
// epoch of the query that marked it, the epoch bumped per query and the slot
// cleared when it wraps.  Bits: one bit per node, the slot's first n / 8
// bytes cleared when the query starts; eight times less memory per slot, so
// the marks of every resident query stay in the device's caches.  A bit is set
// with an atomic OR (neighbours of one step may share a word) and read at
// device scope, so a read after the wave's own OR (which completes in L2)
// cannot be served by a stale first-level cache line.
struct Marks {
  uint8_t* p;
  uint8_t ep;
  bool bits;
  // the raw mark of node t >= 0: a word of 32 bits, or the byte
  __device__ __forceinline__ int raw(int t) const {
    if (bits)
      return (int)__hip_atomic_load(reinterpret_cast<uint32_t*>(p) + (t >> 5), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    return (int)p[t];
  }
  __device__ __forceinline__ bool seen(int raw_mark, int t) const {
    return bits ? (((uint32_t)raw_mark >> (t & 31)) & 1u) != 0u : raw_mark == (int)ep;
  }
  __device__ __forceinline__ bool seen(int t) const { return seen(raw(t), t); }
  __device__ __forceinline__ void mark(int t) const {
    if (bits) atomicOr(reinterpret_cast<uint32_t*>(p) + (t >> 5), 1u << (t & 31));
    else p[t] = ep;
  }
};

template <int FUNC>
__device__ __forceinline__ float term(float q, float x) {
  if constexpr (FUNC == kL2Sqr16 || FUNC == kL2Sqr) {
    const float t = q - x;
    return t * t;
  } else {
    return q * x;
  }
}

template <int FUNC>
__device__ __forceinline__ float finish(float s) {
  if constexpr (FUNC == kNormCos) {
    // max(0, 1 - max(-1, min(1, dot)))   (hnsw.cc:78-81)
    const float c = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
    const float r = 1.0f - c;
    return r < 0.0f ? 0.0f : r;
  } else if constexpr (FUNC == kNegDot) {
    return -s;
  } else {
    return s;
  }
}

// FP32 distance of the query (LDS) to one stored vector, computed by a lane
// pair in the AVX order of the reference; both lanes return the same value.
template <int D, int FUNC>
__device__ __forceinline__ float dist_pair_f32(const float* __restrict__ qs,
                                               const float* __restrict__ x, int h, int drt) {
  const int d = D > 0 ? D : drt;
  const int q16 = d >> 4;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
  for (int c = 0; c < q16; ++c) {
    const float4 x0 = *reinterpret_cast<const float4*>(x + 16 * c + 4 * h);
    const float4 x1 = *reinterpret_cast<const float4*>(x + 16 * c + 8 + 4 * h);
    const float4 q0 = *reinterpret_cast<const float4*>(qs + 16 * c + 4 * h);
    const float4 q1 = *reinterpret_cast<const float4*>(qs + 16 * c + 8 + 4 * h);
    a0 = a0 + term<FUNC>(q0.x, x0.x);
    a1 = a1 + term<FUNC>(q0.y, x0.y);
    a2 = a2 + term<FUNC>(q0.z, x0.z);
    a3 = a3 + term<FUNC>(q0.w, x0.w);
    a0 = a0 + term<FUNC>(q1.x, x1.x);
    a1 = a1 + term<FUNC>(q1.y, x1.y);
    a2 = a2 + term<FUNC>(q1.z, x1.z);
    a3 = a3 + term<FUNC>(q1.w, x1.w);
  }
  const float p0 = __shfl_xor(a0, 1), p1 = __shfl_xor(a1, 1);
  const float p2 = __shfl_xor(a2, 1), p3 = __shfl_xor(a3, 1);
  // accumulators 0..3 live in the even lane, 4..7 in the odd one
  const float l0 = h ? p0 : a0, l1 = h ? p1 : a1, l2 = h ? p2 : a2, l3 = h ? p3 : a3;
  const float u0 = h ? a0 : p0, u1 = h ? a1 : p1, u2 = h ? a2 : p2, u3 = h ? a3 : p3;
  float s;
  if constexpr (FUNC == kL2Sqr16) {
    s = l0 + l1;
    s = s + l2;
    s = s + l3;
    s = s + u0;
    s = s + u1;
    s = s + u2;
    s = s + u3;
  } else {
    float s0 = l0 + u0, s1 = l1 + u1, s2 = l2 + u2, s3 = l3 + u3;
    int c = 16 * q16;
    for (; c + 4 <= d; c += 4) {
      const float4 x4 = *reinterpret_cast<const float4*>(x + c);
      const float4 q4 = *reinterpret_cast<const float4*>(qs + c);
      s0 = s0 + term<FUNC>(q4.x, x4.x);
      s1 = s1 + term<FUNC>(q4.y, x4.y);
      s2 = s2 + term<FUNC>(q4.z, x4.z);
      s3 = s3 + term<FUNC>(q4.w, x4.w);
    }
    s = s0 + s1;
    s = s + s2;
    s = s + s3;
    for (; c < d; ++c) s = s + term<FUNC>(qs[c], x[c]);
  }
  return finish<FUNC>(s);
}

// L1NormSIMD / LInfNormSIMD (distcomp_lp.cc:78-247) by a lane pair: lane h
// keeps SSE lanes 2h and 2h+1, fed by every group of 4 floats in order; the
// four lanes are then summed left to right in float, the tail added in
// double and the result returned as float (L1), or the largest |q - x| taken
// (LInf, independent of order).  Both lanes return the same value.
template <int FUNC>
__device__ __forceinline__ float dist_pair_l1inf(const float* __restrict__ qs,
                                                 const float* __restrict__ x, int h, int d) {
  const int q4 = d >> 2;
  float a0 = 0.f, a1 = 0.f;
  for (int g = 0; g < q4; ++g) {
    const float2 xv = *reinterpret_cast<const float2*>(x + 4 * g + 2 * h);
    const float2 qv = *reinterpret_cast<const float2*>(qs + 4 * g + 2 * h);
    const float t0 = fabsf(qv.x - xv.x), t1 = fabsf(qv.y - xv.y);
    if constexpr (FUNC == kL1) {
      a0 = a0 + t0;
      a1 = a1 + t1;
    } else {
      a0 = a0 < t0 ? t0 : a0;
      a1 = a1 < t1 ? t1 : a1;
    }
  }
  const float p0 = __shfl_xor(a0, 1), p1 = __shfl_xor(a1, 1);
  const float l0 = h ? p0 : a0, l1 = h ? p1 : a1;  // SSE lanes 0, 1
  const float u0 = h ? a0 : p0, u1 = h ? a1 : p1;  // SSE lanes 2, 3
  if constexpr (FUNC == kL1) {
    float s = l0 + l1;
    s = s + u0;
    s = s + u1;
    double r = s;
    for (int c = 4 * q4; c < d; ++c) r += (double)fabsf(qs[c] - x[c]);
    return (float)r;
  } else {
    const float m01 = l0 < l1 ? l1 : l0, m23 = u0 < u1 ? u1 : u0;
    float r = m01 < m23 ? m23 : m01;
    for (int c = 4 * q4; c < d; ++c) {
      const float t = fabsf(qs[c] - x[c]);
      r = r < t ? t : r;
    }
    return r;
  }
}

// The spaces' distances of a regular index (L2SqrSIMD + sqrt, distcomp_lp.cc:
// 304-369; NormScalarProductSIMD, distcomp_scalar.cc:83-168, inside
// CosineSimilarity :267-271; ScalarProductSIMD :194-245) by a lane pair, the
// data point x being the left operand: lane h keeps SSE lanes 2h and 2h+1 of
// every group of 4 floats; the four lanes are summed left to right, then the
// tail is added element by element in float.  Both lanes return the same value.
template <int FUNC>
__device__ __forceinline__ float dist_pair_space(const float* __restrict__ qs,
                                                 const float* __restrict__ x, int h, int d) {
  const int q4 = d >> 2;
  float a0 = 0.f, a1 = 0.f, n0 = 0.f, n1 = 0.f, m0 = 0.f, m1 = 0.f;
  for (int g = 0; g < q4; ++g) {
    const float2 xv = *reinterpret_cast<const float2*>(x + 4 * g + 2 * h);
    const float2 qv = *reinterpret_cast<const float2*>(qs + 4 * g + 2 * h);
    if constexpr (FUNC == kSpaceL2) {
      const float t0 = xv.x - qv.x, t1 = xv.y - qv.y;
      a0 = a0 + t0 * t0;
      a1 = a1 + t1 * t1;
    } else {
      a0 = a0 + xv.x * qv.x;
      a1 = a1 + xv.y * qv.y;
      if constexpr (FUNC == kSpaceCos) {
        n0 = n0 + xv.x * xv.x;
        n1 = n1 + xv.y * xv.y;
        m0 = m0 + qv.x * qv.x;
        m1 = m1 + qv.y * qv.y;
      }
    }
  }
  auto lanes4 = [h](float v0, float v1) {  // SSE lanes 0 + 1 + 2 + 3, left to right
    const float p0 = __shfl_xor(v0, 1), p1 = __shfl_xor(v1, 1);
    const float l0 = h ? p0 : v0, l1 = h ? p1 : v1;
    const float u0 = h ? v0 : p0, u1 = h ? v1 : p1;
    float s = l0 + l1;
    s = s + u0;
    return s + u1;
  };
  float s = lanes4(a0, a1);
  if constexpr (FUNC == kSpaceL2) {
    for (int c = 4 * q4; c < d; ++c) {
      const float t = x[c] - qs[c];
      s = s + t * t;
    }
    return sqrtf(s);
  } else if constexpr (FUNC == kSpaceNegDot) {
    for (int c = 4 * q4; c < d; ++c) s = s + x[c] * qs[c];
    return -s;
  } else {
    float nx = lanes4(n0, n1), nq = lanes4(m0, m1);
    for (int c = 4 * q4; c < d; ++c) {
      s = s + x[c] * qs[c];
      nx = nx + x[c] * x[c];
      nq = nq + qs[c] * qs[c];
    }
    // 0 when either norm is below 2 FLT_MIN, else the clamped normalised
    // product; the distance is max(0, 1 - that)
    const float eps = 2.0f * 1.17549435e-38f;
    float c = 0.0f;
    if (!(nx < eps || nq < eps)) {
      const float r = s / sqrtf(nx) / sqrtf(nq);
      const float lo = r < 1.0f ? r : 1.0f;
      c = -1.0f < lo ? lo : -1.0f;
    }
    const float v = 1.0f - c;
    return 0.0f < v ? v : 0.0f;
  }
}

// FP32 distance in the reference's order for the function
template <int D, int FUNC>
__device__ __forceinline__ float dist_f32(const float* __restrict__ qs, const float* __restrict__ x,
                                          int h, int drt) {
  if constexpr (FUNC == kL1 || FUNC == kLInf) return dist_pair_l1inf<FUNC>(qs, x, h, D > 0 ? D : drt);
  else if constexpr (FUNC >= kSpaceL2) return dist_pair_space<FUNC>(qs, x, h, D > 0 ? D : drt);
  else return dist_pair_f32<D, FUNC>(qs, x, h, drt);
}

__device__ __forceinline__ float bf16_to_f32(uint32_t bits16) {
  return __uint_as_float(bits16 << 16);
}

// one term of a reduced-precision distance, and the lane pair's combination
template <int FUNC>
__device__ __forceinline__ float lp_acc(float acc, float q, float x) {
  if constexpr (FUNC == kL2Sqr16 || FUNC == kL2Sqr) {
    const float t = q - x;
    return __builtin_fmaf(t, t, acc);
  } else if constexpr (FUNC == kL1) {
    return acc + fabsf(q - x);
  } else if constexpr (FUNC == kLInf) {
    const float t = fabsf(q - x);
    return acc < t ? t : acc;
  } else {
    return __builtin_fmaf(q, x, acc);
  }
}
template <int FUNC>
__device__ __forceinline__ float lp_pair(float acc) {
  const float p = __shfl_xor(acc, 1);
  if constexpr (FUNC == kLInf) return acc < p ? p : acc;
  else return acc + p;
}

// BF16 storage: the query is rounded to BF16 too (round to nearest even,
// as the stored vectors are), so a distance is packed products
// (v_dot2_f32_bf16, two coordinates per instruction, FP32 accumulation).
// The rounded query sits right after the FP32 one: word i of qb = coordinates
// 2i and 2i+1, and qb[dstride / 2] = the sum of its squares.
// s_waitcnt vmcnt(0): every vector memory operation of this wave, stores and
// atomics without return included, has been performed (a workgroup-scope
// release fence does not wait for them, and an agent-scope one also writes the
// L2 back, which marks that only this CU reads do not need)
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ const uint32_t* query_bf16(const float* qs, int dstride) {
  return reinterpret_cast<const uint32_t*>(qs + dstride);
}

__device__ __forceinline__ uint32_t bf16_rne(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float dot2_bf16(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, a),
                                         __builtin_bit_cast(bf16x2_t, b), c, false);
}

// Writes the BF16 query, its squared norm and the norm of its rounding error
// |q - q~| (the differences exact in FP32, their squares summed in FP64,
// rounded up: the BF16 screen's bound) after qs[0..dstride); one wave.
__device__ __forceinline__ void make_query_bf16(float* qs, int dstride, int lane) {
  uint32_t* qb = reinterpret_cast<uint32_t*>(qs + dstride);
  float part = 0.f;
  double e2 = 0.0;
  for (int i = lane; i < dstride / 2; i += 64) {
    const uint32_t w = bf16_rne(qs[2 * i]) | (bf16_rne(qs[2 * i + 1]) << 16);
    qb[i] = w;
    part = dot2_bf16(w, w, part);
    const double d0 = (double)qs[2 * i] - (double)__uint_as_float(w << 16);
    const double d1 = (double)qs[2 * i + 1] - (double)__uint_as_float(w & 0xffff0000u);
    e2 += d0 * d0 + d1 * d1;
  }
  // fixed butterfly order: every wave of every kernel gets the same sum
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    part = part + __shfl_xor(part, o);
    e2 += __shfl_xor(e2, o);
  }
  if (lane == 0) {
    reinterpret_cast<float*>(qb)[dstride / 2] = part;
    reinterpret_cast<float*>(qb)[dstride / 2 + 1] = (float)(sqrt(e2) * (1.0 + 0x1p-20)) + 1e-37f;
  }
}

// The BF16 distance from the lane's partial products: x.q over its
// coordinates (sxq) and, for l2, x.x (sxx); the lane pair's halves are added
// and l2 is |x|^2 - 2 x.q + |q|^2 (clamped at 0).  Integer coordinates below
// 256 keep every partial sum an exact integer below 2^24 (d <= 128), so there
// the result equals the FP32 distance.
template <int FUNC>
__device__ __forceinline__ float dot_finish(float sxq, float sxx, const uint32_t* qb, int dstride) {
  sxq = sxq + __shfl_xor(sxq, 1);
  if constexpr (FUNC == kL2Sqr16 || FUNC == kL2Sqr) {
    sxx = sxx + __shfl_xor(sxx, 1);
    const float qq = reinterpret_cast<const float*>(qb)[dstride / 2];
    const float s = (sxx - 2.0f * sxq) + qq;
    return s < 0.0f ? 0.0f : s;
  } else {
    return finish<FUNC>(sxq);
  }
}

// Reduced-precision distance (BF16 or per-vector-scaled FP8 e4m3 storage,
// FP32 accumulation).  A lane pair splits every 16-element chunk into two
// halves; the order is not the reference's (the stored values are not the
// reference's either) -- the exact FP32 rerank restores reference distances.
template <int D, int FUNC, int ST>
__device__ __forceinline__ float dist_pair_lp(const float* __restrict__ qs, const DevIndex& ix,
                                              int64_t node, int h, int drt) {
  const int d = D > 0 ? D : drt;
  const int q16 = d >> 4;
  if constexpr (ST == kBF16 && FUNC <= kNegDot) {
    constexpr bool L2 = FUNC == kL2Sqr16 || FUNC == kL2Sqr;
    const uint32_t* qb = query_bf16(qs, ix.dstride);
    const uint16_t* x = ix.vec16 + node * ix.dstride;
    float sxq = 0.f, sxx = 0.f;
#pragma unroll
    for (int c = 0; c < q16; ++c) {
      const int e0 = 16 * c + 8 * h;
      const uint4 xr = *reinterpret_cast<const uint4*>(x + e0);
      const uint4 qr = *reinterpret_cast<const uint4*>(qb + e0 / 2);
      sxq = dot2_bf16(xr.x, qr.x, sxq);
      sxq = dot2_bf16(xr.y, qr.y, sxq);
      sxq = dot2_bf16(xr.z, qr.z, sxq);
      sxq = dot2_bf16(xr.w, qr.w, sxq);
      if constexpr (L2) {
        sxx = dot2_bf16(xr.x, xr.x, sxx);
        sxx = dot2_bf16(xr.y, xr.y, sxx);
        sxx = dot2_bf16(xr.z, xr.z, sxx);
        sxx = dot2_bf16(xr.w, xr.w, sxx);
      }
    }
    for (int c = 16 * q16 + h; c < d; c += 2) {
      const uint32_t xw = x[c];
      const uint32_t qw = (qb[c >> 1] >> (16 * (c & 1))) & 0xffffu;
      sxq = dot2_bf16(xw, qw, sxq);
      if constexpr (L2) sxx = dot2_bf16(xw, xw, sxx);
    }
    return dot_finish<FUNC>(sxq, sxx, qb, ix.dstride);
  }
  float acc = 0.f;
  float scale = 1.f;
  if constexpr (ST == kFP8) scale = ix.scale8[node];
#pragma unroll
  for (int c = 0; c < q16; ++c) {
    const int e0 = 16 * c + 8 * h;
    float xv[8];
    if constexpr (ST == kBF16) {
      const uint4 raw =
          *reinterpret_cast<const uint4*>(ix.vec16 + node * ix.dstride + e0);
      const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        xv[2 * j] = bf16_to_f32(w[j] & 0xffffu);
        xv[2 * j + 1] = bf16_to_f32(w[j] >> 16);
      }
    } else {
      const uint2 raw = *reinterpret_cast<const uint2*>(ix.vec8 + node * ix.dstride + e0);
      const uint32_t w[2] = {raw.x, raw.y};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[j], false);
        const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[j], true);
        xv[4 * j] = lo[0] * scale;
        xv[4 * j + 1] = lo[1] * scale;
        xv[4 * j + 2] = hi[0] * scale;
        xv[4 * j + 3] = hi[1] * scale;
      }
    }
    const float4 qa = *reinterpret_cast<const float4*>(qs + e0);
    const float4 qb = *reinterpret_cast<const float4*>(qs + e0 + 4);
    const float qv[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) acc = lp_acc<FUNC>(acc, qv[j], xv[j]);
  }
  for (int c = 16 * q16 + h; c < d; c += 2) {
    float xv;
    if constexpr (ST == kBF16) {
      xv = bf16_to_f32(ix.vec16[node * ix.dstride + c]);
    } else {
      const auto p = __builtin_amdgcn_cvt_pk_f32_fp8((int)ix.vec8[node * ix.dstride + c], false);
      xv = p[0] * scale;
    }
    acc = lp_acc<FUNC>(acc, qs[c], xv);
  }
  return finish<FUNC>(lp_pair<FUNC>(acc));
}

template <int D, int FUNC, int ST>
__device__ __forceinline__ float dist_node(const float* __restrict__ qs, const DevIndex& ix,
                                           int64_t node, int h) {
  if constexpr (ST == kF32) {
    return dist_f32<D, FUNC>(qs, ix.vec + node * ix.dstride, h, ix.dim);
  } else {
    return dist_pair_lp<D, FUNC, ST>(qs, ix, node, h, ix.dim);
  }
}

// Batched distances of cand[0..ncand) with every load of a round issued before
// any arithmetic: each lane pair owns NV vectors per round, so a beam step's
// ~100 gathers cost one or two memory round trips instead of one per 32
// vectors.  Registers per lane and vector: 4*(D/16)*... = D/2 floats (FP32),
// D/4 (BF16), D/8 (FP8) -- NV is chosen so the round stays near 128 VGPRs.
// The FP32 arithmetic is dist_pair_f32's (the reference's AVX order).
template <int D, int ST>
struct VecRegs;
template <int D>
struct VecRegs<D, kF32> {
  static constexpr int Q16 = D / 16, R4 = (D % 16) / 4;
  float4 c[2 * Q16 + (R4 > 0 ? R4 : 0) + (2 * Q16 + R4 == 0 ? 1 : 0)];
};
template <int D>
struct VecRegs<D, kBF16> {
  static constexpr int Q16 = D / 16, T = D % 16;
  uint4 c[Q16 > 0 ? Q16 : 1];
  uint16_t t[(T + 1) / 2 > 0 ? (T + 1) / 2 : 1];
};
template <int D>
struct VecRegs<D, kFP8> {
  static constexpr int Q16 = D / 16, T = D % 16;
  uint2 c[Q16 > 0 ? Q16 : 1];
  uint8_t t[(T + 1) / 2 > 0 ? (T + 1) / 2 : 1];
  float scale;
};

template <int D, int ST>
__device__ __forceinline__ void vec_load(VecRegs<D, ST>& r, const DevIndex& ix, int64_t node, int h) {
  constexpr int Q16 = D / 16;
  if constexpr (ST == kF32) {
    const float* x = ix.vec + node * ix.dstride;
#pragma unroll
    for (int c = 0; c < Q16; ++c) {
      r.c[2 * c] = *reinterpret_cast<const float4*>(x + 16 * c + 4 * h);
      r.c[2 * c + 1] = *reinterpret_cast<const float4*>(x + 16 * c + 8 + 4 * h);
    }
#pragma unroll
    for (int t = 0; t < VecRegs<D, kF32>::R4; ++t)
      r.c[2 * Q16 + t] = *reinterpret_cast<const float4*>(x + 16 * Q16 + 4 * t);
  } else if constexpr (ST == kBF16) {
    const uint16_t* x = ix.vec16 + node * ix.dstride;
#pragma unroll
    for (int c = 0; c < Q16; ++c) r.c[c] = *reinterpret_cast<const uint4*>(x + 16 * c + 8 * h);
#pragma unroll
    for (int t = 0; 2 * t + h < D % 16; ++t) r.t[t] = x[16 * Q16 + 2 * t + h];
  } else {
    const uint8_t* x = ix.vec8 + node * ix.dstride;
#pragma unroll
    for (int c = 0; c < Q16; ++c) r.c[c] = *reinterpret_cast<const uint2*>(x + 16 * c + 8 * h);
#pragma unroll
    for (int t = 0; 2 * t + h < D % 16; ++t) r.t[t] = x[16 * Q16 + 2 * t + h];
    r.scale = ix.scale8[node];
  }
}

template <int D, int FUNC, int ST>
__device__ __forceinline__ float vec_dist(const VecRegs<D, ST>& r, const float* __restrict__ qs,
                                          int h, int dstride) {
  constexpr int Q16 = D / 16;
  if constexpr (ST == kBF16 && FUNC <= kNegDot) {
    // the packed products of dist_pair_lp, in the same order
    constexpr bool L2 = FUNC == kL2Sqr16 || FUNC == kL2Sqr;
    const uint32_t* qb = query_bf16(qs, dstride);
    float sxq = 0.f, sxx = 0.f;
#pragma unroll
    for (int c = 0; c < Q16; ++c) {
      const uint4 xr = r.c[c];
      const uint4 qr = *reinterpret_cast<const uint4*>(qb + 8 * c + 4 * h);
      sxq = dot2_bf16(xr.x, qr.x, sxq);
      sxq = dot2_bf16(xr.y, qr.y, sxq);
      sxq = dot2_bf16(xr.z, qr.z, sxq);
      sxq = dot2_bf16(xr.w, qr.w, sxq);
      if constexpr (L2) {
        sxx = dot2_bf16(xr.x, xr.x, sxx);
        sxx = dot2_bf16(xr.y, xr.y, sxx);
        sxx = dot2_bf16(xr.z, xr.z, sxx);
        sxx = dot2_bf16(xr.w, xr.w, sxx);
      }
    }
#pragma unroll
    for (int t = 0; 2 * t + h < D % 16; ++t) {
      const int c = 16 * Q16 + 2 * t + h;
      const uint32_t xw = r.t[t];
      const uint32_t qw = (qb[c >> 1] >> (16 * (c & 1))) & 0xffffu;
      sxq = dot2_bf16(xw, qw, sxq);
      if constexpr (L2) sxx = dot2_bf16(xw, xw, sxx);
    }
    return dot_finish<FUNC>(sxq, sxx, qb, dstride);
  } else if constexpr (ST == kF32) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
    for (int c = 0; c < Q16; ++c) {
      const float4 x0 = r.c[2 * c], x1 = r.c[2 * c + 1];
      const float4 q0 = *reinterpret_cast<const float4*>(qs + 16 * c + 4 * h);
      const float4 q1 = *reinterpret_cast<const float4*>(qs + 16 * c + 8 + 4 * h);
      a0 = a0 + term<FUNC>(q0.x, x0.x);
      a1 = a1 + term<FUNC>(q0.y, x0.y);
      a2 = a2 + term<FUNC>(q0.z, x0.z);
      a3 = a3 + term<FUNC>(q0.w, x0.w);
      a0 = a0 + term<FUNC>(q1.x, x1.x);
      a1 = a1 + term<FUNC>(q1.y, x1.y);
      a2 = a2 + term<FUNC>(q1.z, x1.z);
      a3 = a3 + term<FUNC>(q1.w, x1.w);
    }
    const float p0 = __shfl_xor(a0, 1), p1 = __shfl_xor(a1, 1);
    const float p2 = __shfl_xor(a2, 1), p3 = __shfl_xor(a3, 1);
    const float l0 = h ? p0 : a0, l1 = h ? p1 : a1, l2 = h ? p2 : a2, l3 = h ? p3 : a3;
    const float u0 = h ? a0 : p0, u1 = h ? a1 : p1, u2 = h ? a2 : p2, u3 = h ? a3 : p3;
    float s;
    if constexpr (FUNC == kL2Sqr16) {
      s = l0 + l1;
      s = s + l2;
      s = s + l3;
      s = s + u0;
      s = s + u1;
      s = s + u2;
      s = s + u3;
    } else {
      float s0 = l0 + u0, s1 = l1 + u1, s2 = l2 + u2, s3 = l3 + u3;
#pragma unroll
      for (int t = 0; t < VecRegs<D, kF32>::R4; ++t) {
        const float4 x4 = r.c[2 * Q16 + t];
        const float4 q4 = *reinterpret_cast<const float4*>(qs + 16 * Q16 + 4 * t);
        s0 = s0 + term<FUNC>(q4.x, x4.x);
        s1 = s1 + term<FUNC>(q4.y, x4.y);
        s2 = s2 + term<FUNC>(q4.z, x4.z);
        s3 = s3 + term<FUNC>(q4.w, x4.w);
      }
      s = s0 + s1;
      s = s + s2;
      s = s + s3;
    }
    return finish<FUNC>(s);
  } else {
    float acc = 0.f;
    const float scale = [&] {
      if constexpr (ST == kFP8) return r.scale; else return 1.f;
    }();
#pragma unroll
    for (int c = 0; c < Q16; ++c) {
      const int e0 = 16 * c + 8 * h;
      float xv[8];
      if constexpr (ST == kBF16) {
        const uint32_t wv[4] = {r.c[c].x, r.c[c].y, r.c[c].z, r.c[c].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          xv[2 * j] = bf16_to_f32(wv[j] & 0xffffu);
          xv[2 * j + 1] = bf16_to_f32(wv[j] >> 16);
        }
      } else {
        const uint32_t wv[2] = {r.c[c].x, r.c[c].y};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)wv[j], false);
          const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)wv[j], true);
          xv[4 * j] = lo[0] * scale;
          xv[4 * j + 1] = lo[1] * scale;
          xv[4 * j + 2] = hi[0] * scale;
          xv[4 * j + 3] = hi[1] * scale;
        }
      }
      const float4 qa = *reinterpret_cast<const float4*>(qs + e0);
      const float4 qb = *reinterpret_cast<const float4*>(qs + e0 + 4);
      const float qv[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
#pragma unroll
      for (int j = 0; j < 8; ++j) acc = lp_acc<FUNC>(acc, qv[j], xv[j]);
    }
#pragma unroll
    for (int t = 0; 2 * t + h < D % 16; ++t) {
      const int c = 16 * Q16 + 2 * t + h;
      float xv;
      if constexpr (ST == kBF16) {
        xv = bf16_to_f32(r.t[t]);
      } else {
        xv = __builtin_amdgcn_cvt_pk_f32_fp8((int)r.t[t], false)[0] * scale;
      }
      acc = lp_acc<FUNC>(acc, qs[c], xv);
    }
    return finish<FUNC>(lp_pair<FUNC>(acc));
  }
}

template <int ST>
struct WideNV {  // vectors in flight per lane pair: about 128 registers of payload
  static constexpr int value = ST == kF32 ? 2 : (ST == kBF16 ? 5 : 8);
};

// Vectors per lane pair and round of dist_cands: PATH 0 = one (dist_node),
// 1 = WideNV (a wave with the register file of one wave per SIMD), 2 = two
// BF16 / four FP8 vectors (two waves per SIMD; FP32 keeps one), 3 = one, its
// loads all issued before any arithmetic (the exact search).
template <int ST, int PATH>
struct PathNV {
  static constexpr int value =
      PATH == 1 ? WideNV<ST>::value
                : (PATH == 2 ? (ST == kF32 ? 0 : (ST == kBF16 ? 2 : 4)) : (PATH == 3 ? 1 : 0));
};

// cdist[s] = distance of cand[s] for s < ncand (fallback: a valid node id used
// by idle lane pairs so every shuffle sees all lanes).  The ids come back from
// LDS, so each passes checked_node before it addresses a vector.
template <int D, int FUNC, int ST, int WIDE>
__device__ __forceinline__ void dist_cands(const float* __restrict__ qs, const DevIndex& ix,
                                           const int* cand, float* cdist, int ncand, int lane,
                                           int h, int fallback, int* err) {
  constexpr int NV = PathNV<ST, WIDE>::value;
  if constexpr (NV > 0 && D > 0 && D % 4 == 0 && (ST != kF32 || FUNC <= kNegDot)) {
    for (int b = 0; b < ncand; b += 32 * NV) {
      VecRegs<D, ST> r[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int s = b + 32 * v + (lane >> 1);
        const int id = checked_node(s < ncand ? cand[s] : fallback, (int)ix.n, err);
        vec_load<D, ST>(r[v], ix, id, h);
      }
      // keep the scheduler from interleaving the loads with the arithmetic
      // (it would otherwise wait on each pair of loads in turn)
      if constexpr (WIDE == 3) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int s = b + 32 * v + (lane >> 1);
        const float dd = vec_dist<D, FUNC, ST>(r[v], qs, h, ix.dstride);
        if (s < ncand && h == 0) cdist[s] = dd;
      }
    }
  } else {
    for (int b = 0; b < ncand; b += 32) {
      const int s = b + (lane >> 1);
      const bool ok = s < ncand;
      const int id = checked_node(ok ? cand[s] : fallback, (int)ix.n, err);
      const float dd = dist_node<D, FUNC, ST>(qs, ix, id, h);
      if (ok && h == 0) cdist[s] = dd;
    }
  }
}

// Candidates dist_cands() covers per round of loads (one candidate per lane
// pair, or NV per pair on the wide path).
template <int D, int FUNC, int ST, int WIDE>
struct DistRound {
  static constexpr int NV = PathNV<ST, WIDE>::value;
  static constexpr int value =
      (NV > 0 && D > 0 && D % 4 == 0 && (ST != kF32 || FUNC <= kNegDot)) ? 32 * NV : 32;
};

// Per-wave LDS: the sorted array, the query, two scratch regions of 8 * ent
// bytes (ent = neighbour entries per step, a multiple of 64, >= 128) whose
// uses do not overlap in time within a step, and the merge's position bitmap:
//   region X: dedupe table (2 * ent slots) -> fresh neighbours + distances
//             -> sorted buffer of the exact search
//   region Y: picks of the beam step -> filtered items -> sorted items
struct WaveLds {
  uint2* A;        // [cap] sorted array: x = key bits, y = internal id | kUsedBit
  float* qs;       // [dstride] query, then [dstride / 2 + 1] its BF16 form (query_bf16)
  uint32_t* htab;  // X [2 * ent] per-step dedupe table of the beam search
  int* cand;       // X [ent] fresh neighbours of this step (list order)
  float* cdist;    // X [ent] their distances
  float* skey;     // X [ent] buffer (sorted)
  int* sid;        // X
  int* picks;      // Y [beam] nodes expanded by a beam step
  float* bkey;     // Y [ent] buffer (list order)
  int* bid;        // Y
  uint2* items;    // Y [ent] beam step: admitted items, then the same sorted
  // beam mode only (not allocated for the other modes):
  uint4* meta;     // [cap / 64 + 2] beam merge: per chunk, which positions new
                   // items take (x, y) and how many land below it (z)
  uint2* pfa;      // [24] beam step: first unexpanded items of the array
  int* pfpos;      // [24] their positions
  uint2* pfq;      // [16] the next step's picks, claimed in advance
  int* ctl;        // [24] two-wave queries: two alternating 8-word commands
                   // from the first wave to the second, and the second
                   // wave's count of admitted items (word 16)
  uint2* items2;   // [ent] two-wave queries with a deferred rebuild: the items
                   // buffer of every other step (the launch allocates it)
};

__device__ __forceinline__ WaveLds carve(unsigned char* base, int cap, int dstride, int ent) {
  WaveLds w;
  w.A = reinterpret_cast<uint2*>(base);
  w.qs = reinterpret_cast<float*>(w.A + ((cap + 1) & ~1));  // keeps 16-byte alignment
  int* x = reinterpret_cast<int*>(w.qs + 2 * dstride);  // query area: see query_bf16
  int* y = x + 2 * ent;
  w.htab = reinterpret_cast<uint32_t*>(x);
  w.cand = x;
  w.cdist = reinterpret_cast<float*>(x + ent);
  w.skey = reinterpret_cast<float*>(x);
  w.sid = x + ent;
  w.picks = y;
  w.bkey = reinterpret_cast<float*>(y);
  w.bid = y + ent;
  w.items = reinterpret_cast<uint2*>(y);
  w.meta = reinterpret_cast<uint4*>(y + 2 * ent);
  w.pfa = reinterpret_cast<uint2*>(w.meta + (cap / 64 + 2));
  w.pfq = w.pfa + 24;
  w.pfpos = reinterpret_cast<int*>(w.pfq + 16);
  w.ctl = w.pfpos + 24;
  w.items2 = reinterpret_cast<uint2*>(w.ctl + 24);
  return w;
}

__device__ __forceinline__ float akey(const uint2* A, int i) { return __uint_as_float(A[i].x); }

// The candidates cand[s0..s1) the beam step's snapshot admits (distance below
// the last item's, or the array not yet full), as (distance, id) items in
// candidate order at out[0..); returns their number.
__device__ __forceinline__ int admit(const WaveLds& w, int s0, int s1, float topk, int nsnap, int ef,
                                     int lane, uint64_t ltm, uint2* out) {
  int m = 0;
  for (int b = s0; b < s1; b += 64) {
    const int s = b + lane;
    const float dj = s < s1 ? w.cdist[s] : 0.f;
    const bool keep = s < s1 && (dj < topk || nsnap < ef);
    const uint64_t km = __ballot(keep);
    if (keep) out[m + __popcll(km & ltm)] = make_uint2(__float_as_uint(dj), (uint32_t)w.cand[s]);
    m += __popcll(km);
  }
  return m;
}

// number of keys of A[0..n) below v (lower) or not above v (upper)
__device__ __forceinline__ int lower_count(const uint2* A, int n, float v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (akey(A, mid) < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int upper_count(const uint2* A, int n, float v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (akey(A, mid) <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Interleaves cnt new items into A[0..na): item j (lane j) lands at pos_j + j
// where pos_j, nondecreasing in j, is the number of A items before it; A item
// i moves up by #{j : pos_j <= i}.  Anything landing at or beyond C drops off.
// Chunks of 64 items move top-down: a chunk's writes land at or above its own
// base while the next chunk reads below it, so only compiler ordering is
// needed between them (one wave, LDS operations execute in order).
__device__ __forceinline__ void place_items(uint2* A, int na, int C, int lane, int cnt, int pos_j,
                                            uint2 item_j) {
  if (cnt <= 0) return;
  const int lo = __shfl(pos_j, 0);
  const int lo_chunk = lo & ~63;
  auto shift_of = [&](int i, int base) {
    int sh = __popcll(__ballot(lane < cnt && pos_j < base));
    uint64_t inm = __ballot(lane < cnt && pos_j >= base && pos_j < base + 64);
    while (inm) {
      const int j = __ffsll((long long)inm) - 1;
      inm &= inm - 1;
      sh += i >= __builtin_amdgcn_readlane(pos_j, j) ? 1 : 0;
    }
    return sh;
  };
  // two chunks per round: both are read before either is written
  for (int base = ((na - 1) >> 6) << 6; base >= lo_chunk; base -= 128) {
    const int b1 = base - 64;
    const int i0 = base + lane, i1 = b1 + lane;
    const bool mv0 = i0 >= lo && i0 < na;
    const bool mv1 = b1 >= lo_chunk && i1 >= lo && i1 < na;
    uint2 v0 = make_uint2(0u, 0u), v1 = make_uint2(0u, 0u);
    if (mv0) v0 = A[i0];
    if (mv1) v1 = A[i1];
    const int sh0 = shift_of(i0, base);
    const int sh1 = b1 >= lo_chunk ? shift_of(i1, b1) : 0;
    __builtin_amdgcn_wave_barrier();
    if (mv0 && i0 + sh0 < C) A[i0 + sh0] = v0;
    if (mv1 && i1 + sh1 < C) A[i1 + sh1] = v1;
    __builtin_amdgcn_wave_barrier();
  }
  if (lane < cnt && pos_j + lane < C) A[pos_j + lane] = item_j;
}

// Greedy descent over levels maxlevel..1 (hnsw_distfunc_opt.cc:168-198).  A
// full scan of the list is equivalent to taking the first position holding
// the list's minimum when that minimum is below the current distance.
template <int D, int FUNC, int ST>
__device__ __forceinline__ void greedy_descent(const DevIndex& ix, const float* qs, int lane,
                                               int h, int& cur, float& curd,
                                               unsigned long long& nd, unsigned long long& nl) {
  for (int lev = ix.maxlevel; lev > 0; --lev) {
    for (;;) {
      const int off = ix.up_off[cur];
      // cur has a list at level lev: the entry point reaches maxlevel and the
      // loader checks that every level-L neighbour has L levels of its own
      if (off < 0) break;
      const int32_t* lst = ix.up + off + (ix.maxM + 1) * (lev - 1);
      const int cnt = min(lst[0], ix.maxM);
      nd += cnt;
      nl += 1 + cnt;
      float bestd = kInf();
      int bestp = 0x7fffffff;
      for (int b = 0; b < cnt; b += 32) {
        const int s = b + (lane >> 1);
        const bool ok = s < cnt;
        const int t = ok ? lst[1 + s] : cur;
        const float dd = dist_node<D, FUNC, ST>(qs, ix, t, h);
        float kd = (ok && h == 0) ? dd : kInf();
        int kp = (ok && h == 0) ? s : 0x7fffffff;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          const float od = __shfl_xor(kd, o);
          const int op = __shfl_xor(kp, o);
          if (od < kd || (od == kd && op < kp)) {
            kd = od;
            kp = op;
          }
        }
        if (kd < bestd) {  // earlier passes win ties (smaller positions)
          bestd = kd;
          bestp = kp;
        }
      }
      if (bestd < curd) {
        curd = bestd;
        cur = lst[1 + bestp];
      } else {
        break;
      }
    }
  }
}

// position of the first item at or after p that is not expanded yet (n if
// none): one chunk of 64, then four chunks per round trip
__device__ __forceinline__ int first_unused(const uint2* A, int p, int n, int lane) {
  if (p < n) {
    const int i = p + lane;
    const uint64_t um = __ballot(i < n && !(A[i].y & kUsedBit));
    if (um) return p + __ffsll((long long)um) - 1;
    p += 64;
  }
  while (p < n) {
    uint32_t y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = p + 64 * u + lane;
      y[u] = i < n ? A[i].y : kUsedBit;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t um = __ballot(!(y[u] & kUsedBit));
      if (um) return p + 64 * u + __ffsll((long long)um) - 1;
    }
    p += 256;
  }
  return n;
}

struct QueryCounters {
  unsigned long long nexp, nd0, ndu, nl, nties;
};

// Optional FP32 rerank, then the first min(k, n) items of the sorted array in
// (distance, internal id) order, and the query's counters.
template <int D, int FUNC, int ST>
__device__ __forceinline__ void finish_query(const SearchArgs& a, const WaveLds& w, int n, int qi, int k,
                             int lane, int h, const QueryCounters& c) {
  const DevIndex& ix = a.ix;
  uint2* A = w.A;
  if constexpr (ST != kF32) {
    // exact FP32 rerank of the best R candidates of the reduced-precision
    // traversal: distances recomputed from the FP32 vectors in the reference's
    // summation order, then the R items re-sorted by (distance, internal id)
    if (a.rerank > 0) {
      const int R = min(n, max(k, a.rerank * k));
      for (int b = 0; b < R; b += 32) {
        const int s = b + (lane >> 1);
        const bool ok = s < R;
        const int id = ok ? checked_node((int)(A[s].y & ~kUsedBit), (int)ix.n, a.err) : 0;
        const float dd = dist_f32<D, FUNC>(w.qs, ix.vec + (int64_t)id * ix.dstride, h, ix.dim);
        if (ok && h == 0) A[s] = make_uint2(__float_as_uint(dd), (uint32_t)id);
      }
      WSYNC();
      // rank sort, staged in A[R, 2R) (the capacity holds R more items)
      for (int b = 0; b < R; b += 64) {
        const int p = b + lane;
        if (p < R) {
          const uint2 v = A[p];
          const float key = __uint_as_float(v.x);
          int rank = 0;
          for (int j = 0; j < R; ++j) {
            const uint2 o = A[j];
            const float kj = __uint_as_float(o.x);
            rank += (kj < key) || (kj == key && o.y < v.y);
          }
          A[R + rank] = v;
        }
        WSYNC();
      }
      for (int p = lane; p < R; p += 64) A[p] = A[R + p];
      WSYNC();
      n = R;
    }
  }

  // report the first min(k, n) items, ties by internal index
  const int kk = min(k, n);
  const int64_t row = (int64_t)qi * a.out_stride;
  for (int b = 0; b < k; b += 64) {
    const int p = b + lane;
    if (p < kk) {
      const uint2 v = A[p];
      const float key = __uint_as_float(v.x);
      const uint32_t id = v.y & ~kUsedBit;
      const int start = lower_count(A, kk, key);
      int rank = start;
      for (int j = start; j < kk && akey(A, j) == key; ++j)
        rank += ((A[j].y & ~kUsedBit) < id) ? 1 : 0;
      if (rank >= kk) {  // only an unsorted (corrupted) array gets here
        if (a.err) atomicOr(a.err, 2);
        rank = p;
      }
      a.out_ids[row + rank] = ix.ext_id[checked_node((int)id, (int)ix.n, a.err)];
      a.out_dists[row + rank] = key;
      if (a.out_internal) a.out_internal[row + rank] = (int32_t)id;
    } else if (p < k) {
      a.out_ids[row + p] = -1;
      a.out_dists[row + p] = 0.f;
      if (a.out_internal) a.out_internal[row + p] = -1;
    }
  }
  if (lane == 0) {
    a.out_counts[qi] = kk;
    if (a.stats) {
      atomicAdd(&a.stats->expansions, c.nexp);
      atomicAdd(&a.stats->dist_level0, c.nd0);
      atomicAdd(&a.stats->dist_upper, c.ndu);
      atomicAdd(&a.stats->list_ints_read, c.nl);
      atomicAdd(&a.stats->tie_paths, c.nties);
    }
  }
}

// ---------------------------------------------------------------------------
// SearchV1Merge at level 0 for one query held by one wave.
//
// The reference's sorted array (SortArrBI, capacity max(ef, k)) is held in two
// parts whose interleaving is the array: A, the bulk, sorted in LDS, and S, up
// to 64 recently inserted items in registers (lane t holds S[t] and sab, the
// number of A items before it; S[t] precedes A[i] iff sab <= i).  Insertions
// go into S, so an expansion moves no LDS items; S is merged into A in one pass
// only when the next insertions would not fit, and at the end.  Positions,
// the last item, truncation at the capacity and the first unexpanded item are
// computed over the interleaving.
template <int D, int FUNC, int ST>
__device__ __forceinline__ void v1merge_query(const SearchArgs& a, const WaveLds& w, const Marks& vis,
                              int qi, int k, int ef, int lane, int h, uint64_t ltm) {
  const DevIndex& ix = a.ix;
  const int C = max(ef, k);
  uint2* A = w.A;
  unsigned long long nexp = 0, nd0 = 0, ndu = 1, nl = 0, nties = 0;

  int cur = ix.enter;
  float curd = dist_node<D, FUNC, ST>(w.qs, ix, cur, h);
  greedy_descent<D, FUNC, ST>(ix, w.qs, lane, h, cur, curd, ndu, nl);

  if (lane == 0) {
    A[0] = make_uint2(__float_as_uint(curd), (uint32_t)cur);
    vis.mark(cur);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  int na = 1, ns = 0;  // items in A and in S
  float s_key = 0.f;   // lane t < ns: S[t]'s key,
  uint32_t s_id = 0;   //   internal id | kUsedBit once expanded,
  int s_ab = 0;        //   and the number of A items before it
  int ceA = 0;         // every A item below ceA is expanded
  int pf_node = -1, pf_t = -1;  // adjacency row of the node predicted to expand next
  int last_node = -1;
  // per-phase device time (100 MHz ticks) when counters are collected: pick +
  // adjacency, visited, distances, admission + prediction, sort, rank search,
  // insertion + merge of S into A, advance
  const bool timed = a.stats != nullptr;
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t tmark = timed ? __builtin_amdgcn_s_memrealtime() : 0;
  auto lap = [&](int p) __attribute__((always_inline)) {
    if (timed) {
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      ph[p] += now - tmark;
      tmark = now;
    }
  };
  auto rl_f = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
  auto rl_i = [](int v, int l) { return __builtin_amdgcn_readlane(v, l); };
  // merges S into A: one pass moving A items up (place_items); the cursor
  // moves to the new index of the first item that may be unexpanded
  auto flush = [&]() __attribute__((always_inline)) {
    if (ns == 0) return;
    const int c0 = min(ceA, na);
    int nc = c0 + __popcll(__ballot(lane < ns && s_ab <= c0));
    const uint64_t su = __ballot(lane < ns && !(s_id & kUsedBit));
    if (su) {
      const int t = __ffsll((long long)su) - 1;
      nc = min(nc, rl_i(s_ab, t) + t);
    }
    WSYNC();
    place_items(A, na, C, lane, ns, s_ab, make_uint2(__float_as_uint(s_key), s_id));
    na += ns;
    ns = 0;
    ceA = nc;
    WSYNC();
  };

  for (;;) {
    WSYNC();
    // the first unexpanded item (the reference's currElem after its advance,
    // hnsw_distfunc_opt.cc:272-273): the first one of A at or after ceA or the
    // first one of S, whichever comes first in the array
    ceA = first_unused(A, ceA, na, lane);
    const uint64_t su = __ballot(lane < ns && !(s_id & kUsedBit));
    const int ts = su ? __ffsll((long long)su) - 1 : ns;
    const int abt = rl_i(s_ab, min(ts, 63));
    const bool from_a = ceA < na && (ts == ns || ceA < abt);
    const int n = na + ns;
    const int vpos = from_a ? ceA + __popcll(__ballot(lane < ns && s_ab <= ceA)) : (ts < ns ? abt + ts : n);
    if (vpos >= min(n, ef)) break;
    if (nexp >= (unsigned long long)ix.n) {  // every expansion takes a new node
      if (a.err) atomicOr(a.err, 2);
      break;
    }
    int node;
    if (from_a) {
      const uint2 e = A[ceA];
      node = (int)(e.y & ~kUsedBit);
      if (lane == 0) A[ceA].y = e.y | kUsedBit;
      ++ceA;
    } else {
      node = (int)((uint32_t)rl_i((int)s_id, ts) & ~kUsedBit);
      if (lane == ts) s_id |= kUsedBit;
    }
    node = checked_node(node, (int)ix.n, a.err);
    if (node == last_node) {  // an item expanded twice: the state is corrupted
      if (a.err) atomicOr(a.err, 2);
      break;
    }
    last_node = node;
    ++nexp;
    // the last item's key (top_key)
    const bool last_s = ns > 0 && rl_i(s_ab, max(ns - 1, 0)) == na;
    const float topk = last_s ? rl_f(s_key, max(ns - 1, 0)) : (na > 0 ? akey(A, na - 1) : kInf());
    const int nsnap = n;

    // adjacency list (one 128-byte row for maxM0 <= 32), usually prefetched
    int t;
    if (node == pf_node) {
      t = pf_t;
    } else {
      const int32_t* lst = ix.l0 + (int64_t)node * ix.l0stride;
      t = lane < ix.maxM0 ? lst[lane] : -1;
    }
    const bool valid = t >= 0;
    nl += 1 + __popcll(__ballot(valid));
    lap(0);
    bool fresh = false;
    if (valid) fresh = !vis.seen(t);
    if (ix.has_dups) {
      w.cand[lane] = t;
      WSYNC();
      for (int j = 0; j < lane; ++j)
        if (w.cand[j] == t) fresh = false;
      WSYNC();
    }
    if (fresh) vis.mark(t);
    const uint64_t fm = __ballot(fresh);
    const int ncand = __popcll(fm);
    if (fresh) w.cand[__popcll(fm & ltm)] = t;
    nd0 += ncand;
    WSYNC();
    lap(1);
    dist_cands<D, FUNC, ST, 3>(w.qs, ix, w.cand, w.cdist, ncand, lane, h, node, a.err);
    WSYNC();
    lap(2);
    const float dj = lane < ncand ? w.cdist[lane] : 0.f;
    const int cj = lane < ncand ? w.cand[lane] : 0;
    const bool keep = lane < ncand && (dj < topk || nsnap < ef);
    const uint64_t km = __ballot(keep);
    const int m = __popcll(km);

    // The node expanded next is the first unexpanded item, unless a new item
    // sorts before it: start loading its adjacency row now so the load
    // overlaps the sort and insertion below.
    {
      ceA = first_unused(A, ceA, na, lane);  // A changes only in flush, which resets ceA
      const uint64_t su2 = __ballot(lane < ns && !(s_id & kUsedBit));
      const int ts2 = su2 ? __ffsll((long long)su2) - 1 : ns;
      const bool a2 = ceA < na && (ts2 == ns || ceA < rl_i(s_ab, min(ts2, 63)));
      float nk = kInf();
      int nid = -1;
      if (a2) {
        const uint2 e2 = A[ceA];
        nk = __uint_as_float(e2.x);
        nid = (int)(e2.y & ~kUsedBit);
      } else if (ts2 < ns) {
        nk = rl_f(s_key, ts2);
        nid = (int)((uint32_t)rl_i((int)s_id, ts2) & ~kUsedBit);
      }
      float bk = keep ? dj : kInf();
      int bi = keep ? cj : -1;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        const float ok = __shfl_xor(bk, o);
        const int oi = __shfl_xor(bi, o);
        if (ok < bk) {
          bk = ok;
          bi = oi;
        }
      }
      const int guess = (bi >= 0 && bk < nk) ? bi : nid;
      if (guess >= 0) {
        pf_node = guess;
        pf_t = lane < ix.maxM0 ? ix.l0[(int64_t)checked_node(guess, (int)ix.n, a.err) * ix.l0stride + lane]
                               : -1;
      } else {
        pf_node = -1;
      }
    }
    lap(3);

    if (m > 0) {
      if (keep) {
        const int r = __popcll(km & ltm);
        w.bkey[r] = dj;
        w.bid[r] = cj;
      }
      WSYNC();
      // the reference sorts the admitted items with std::sort: stable ranks;
      // an equal pair among more than 16 items needs the exact std::sort
      // order (insertion sort is stable up to 16 items)
      const float mk = lane < m ? w.bkey[lane] : 0.f;
      const int mid = lane < m ? w.bid[lane] : 0;
      int rank = 0;
      bool tie = false;
      for (int j = 0; j < m; ++j) {  // uniform loop: keys broadcast by readlane
        const float kj = rl_f(mk, j);
        rank += (kj < mk) || (kj == mk && j < lane);
        tie |= (kj == mk) && (j != lane);
      }
      tie = tie && lane < m;
      const bool anytie = __ballot(tie) != 0;
      if (anytie && m > 16) {
        if (lane < m) {
          w.skey[lane] = mk;
          w.sid[lane] = mid;
        }
        WSYNC();
        if (lane == 0) emu_std_sort(w.skey, w.sid, m);
      } else if (lane < m) {
        w.skey[rank] = mk;
        w.sid[rank] = mid;
      }
      WSYNC();
      const float bj = lane < m ? w.skey[lane] : 0.f;
      const int bidj = lane < m ? w.sid[lane] : 0;
      lap(4);
      if (ns + m > 64) flush();
      lap(6);
      // items of A below bj (Lj) and not above it (Uj); the same over S
      const int Lj = lane < m ? lower_count(A, na, bj) : 0;
      int Uj = Lj;
      if (lane < m)
        while (Uj < na && akey(A, Uj) == bj) ++Uj;
      int LSj = 0, USj = 0, s_up = 0;  // s_up: new items below S[lane]
      for (int j = 0; j < m; ++j) {
        const float b = rl_f(bj, j);
        const int ls = __popcll(__ballot(lane < ns && s_key < b));
        const int us = __popcll(__ballot(lane < ns && s_key <= b));
        if (lane == j) {
          LSj = ls;
          USj = us;
        }
        s_up += b < s_key ? 1 : 0;
      }
      const bool slow = anytie || (__ballot(lane < m && (Lj != Uj || LSj != USj)) != 0);
      lap(5);
      if (!slow) {
        // distinct keys: sequential insertion == merge truncated to the
        // capacity.  New item j takes S position LSj + j with Lj A items
        // before it; S[t] moves up by the new items below it.
        float* xk = w.skey;
        int* xi = w.sid;
        int* xa = w.bid;
        WSYNC();
        if (lane < ns) {
          const int d = lane + s_up;
          xk[d] = s_key;
          xi[d] = (int)s_id;
          xa[d] = s_ab;
        }
        if (lane < m) {
          const int d = LSj + lane;
          xk[d] = bj;
          xi[d] = bidj;
          xa[d] = Lj;
        }
        WSYNC();
        ns += m;
        if (lane < ns) {
          s_key = xk[lane];
          s_id = (uint32_t)xi[lane];
          s_ab = xa[lane];
        }
        // the last items fall off
        while (na + ns > C) {
          if (ns > 0 && rl_i(s_ab, ns - 1) == na) --ns; else --na;
        }
      } else {
        // Equal keys: replay the reference's item-by-item insertion
        // (push_or_replace_non_empty_exp, sort_arr_bi.h:159-199) over the
        // interleaving: counts over A are precomputed, counts over S are
        // ballots, and S takes the items.
        ++nties;
        for (int j = 0; j < m; ++j) {
          const float b = rl_f(bj, j);
          const int bi = rl_i(bidj, j);
          const int LA = min(rl_i(Lj, j), na);
          const int UA = min(rl_i(Uj, j), na);
          const int nv = na + ns;
          const bool last_is_s = ns > 0 && rl_i(s_ab, max(ns - 1, 0)) == na;
          const float lastkey = last_is_s ? rl_f(s_key, max(ns - 1, 0)) : (na > 0 ? akey(A, na - 1) : kInf());
          if (lastkey <= b) {
            // append after an equal last key while there is room
            if (nv < C) {
              if (lane == ns) {
                s_key = b;
                s_ab = na;
                s_id = (uint32_t)bi;
              }
              ++ns;
            }
          } else {
            const int L = LA + __popcll(__ballot(lane < ns && s_key < b));
            const int U = UA + __popcll(__ballot(lane < ns && s_key <= b));
            // exponential walk down from the last item (sort_arr_bi.h:170-181),
            // then the lower bound inside the last step's window
            int c = nv - 1, prev = c, step = 1;
            while (c > 0 && c >= U) {
              prev = c;
              c -= step;
              step *= 2;
              if (step > c) step = c;
            }
            const int pos = c < prev ? max(c, L) : c;
            if (nv == C) {  // the last item falls off
              if (last_is_s) --ns; else --na;
            }
            const int t0 = __popcll(__ballot(lane < ns && s_ab + lane < pos));
            const float uk = __shfl_up(s_key, 1);
            const int ua = __shfl_up(s_ab, 1);
            const uint32_t ui = (uint32_t)__shfl_up((int)s_id, 1);
            if (lane > t0 && lane <= ns) {
              s_key = uk;
              s_ab = ua;
              s_id = ui;
            }
            if (lane == t0) {
              s_key = b;
              s_ab = pos - t0;
              s_id = (uint32_t)bi;
            }
            ++ns;
          }
        }
      }
      lap(6);
    }
    // the marks of this expansion must land before the next one reads them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    lap(7);
  }
  flush();
  if (timed && lane == 0)
    for (int p = 0; p < 8; ++p) atomicAdd(&a.stats->phase_ticks[p], ph[p]);

  const QueryCounters c{nexp, nd0, ndu, nl, nties};
  finish_query<D, FUNC, ST>(a, w, na, qi, k, lane, h, c);
}


// ---------------------------------------------------------------------------
// Beam mode (an extension; no reference counterpart): each step expands the
// first `beam` unexpanded items of the sorted array together, so one round of
// adjacency, visited and vector loads serves several expansions and the
// serial chain of a query shortens by that factor.  Items are ordered by
// (distance, internal id), which makes every key distinct, so new items always
// take the parallel merge.  The array, its capacity max(ef, k) and the stop
// rule are V1Merge's.
__device__ __forceinline__ bool kid_less(float ka, uint32_t ia, float kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__device__ __forceinline__ int lower_count_kid(const uint2* A, int n, float v, uint32_t id) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const uint2 e = A[mid];
    if (kid_less(__uint_as_float(e.x), e.y & ~kUsedBit, v, id)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// The value lane (lane ^ d) holds.  d = 1, 2: DPP quad permutes; 4, 8: DPP
// row shifts (row_shl:d brings lane i + d, row_shr:d lane i - d, within rows
// of 16); 16, 32: v_permlane16_swap / v_permlane32_swap with both operands v,
// which return {v with its odd rows (upper half) replaced by the even rows
// (lower half), v with its even rows (lower half) replaced by the odd rows
// (upper half)}.  d is a constant after unrolling.
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int d, int lane) {
  switch (d) {
    case 1: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    case 4: {
      const int up = __builtin_amdgcn_update_dpp(0, (int)v, 0x104, 0xF, 0xF, false);
      const int dn = __builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
      return (uint32_t)((lane & 4) ? dn : up);
    }
    case 8: {
      const int up = __builtin_amdgcn_update_dpp(0, (int)v, 0x108, 0xF, 0xF, false);
      const int dn = __builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
      return (uint32_t)((lane & 8) ? dn : up);
    }
    case 16: {
      const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
      return (lane & 16) ? r[0] : r[1];
    }
    default: {
      const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
      return (lane & 32) ? r[0] : r[1];
    }
  }
}

// Sorts the 64 * R items held as item r * 64 + lane = (k[r], id[r]) into
// ascending (key, id) order with Batcher's bitonic network: partners 64 or
// more apart sit in the same lane, nearer ones are exchanged across lanes.
template <int R>
__device__ __forceinline__ void bitonic_sort(float (&k)[R], uint32_t (&id)[R], int lane) {
#pragma unroll
  for (int s = 2; s <= 64 * R; s <<= 1) {
#pragma unroll
    for (int d = s >> 1; d > 0; d >>= 1) {
      if (d >= 64) {
        const int dr = d >> 6;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & dr) == 0) {
            const int r2 = r | dr;
            const bool asc = ((r << 6) & s) == 0;
            const bool sw = asc ? kid_less(k[r2], id[r2], k[r], id[r])
                                : kid_less(k[r], id[r], k[r2], id[r2]);
            if (sw) {
              const float tk = k[r];
              k[r] = k[r2];
              k[r2] = tk;
              const uint32_t ti = id[r];
              id[r] = id[r2];
              id[r2] = ti;
            }
          }
        }
      } else {
        const bool lower = (lane & d) == 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool asc = (((r << 6) | lane) & s) == 0;
          const float pk = __uint_as_float(xor_lane(__float_as_uint(k[r]), d, lane));
          const uint32_t pi = xor_lane(id[r], d, lane);
          const bool pl = kid_less(pk, pi, k[r], id[r]);
          // the lower slot keeps the smaller item when ascending
          if (lower == asc ? pl : !pl) {
            k[r] = pk;
            id[r] = pi;
          }
        }
      }
    }
  }
}

// Two waves per query (kModeBeamPair): the first wave runs the search and
// hands the second part of each step through commands in LDS.  Posting a
// command is one workgroup barrier: the second wave waits at it, then reads
// the command and carries it out; the first wave posts an empty command when
// it needs the second one's results.  Commands alternate between two 8-word
// slots, so the next one can be written while the last is still being read.
// Only post() and walk_pair() pass barriers, and both waves derive the number
// walk_pair() passes from the same command, so the barrier counts match.
enum PairCmd : int { kCmdNop = 0, kCmdDist = 1, kCmdWalk = 2, kCmdRank = 3, kCmdExit = 4 };

// The barrier between the two waves of a query.  Every LDS access this wave
// issued completes before it arrives (s_waitcnt lgkmcnt(0), written out):
// the release fence alone left a hand-off write (the second wave's admitted
// count, ds_write then s_barrier on the loop's back edge) in flight at the
// barrier in the generated code, and with other blocks' LDS traffic on the CU
// the first wave then read the word, and the items, before the write landed.
__device__ __forceinline__ void pair_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt unconstrained
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

struct PairLink {
  int* ctl;            // the command slots (16 words) and the second wave's count
  int seq;             // commands posted so far
  const uint2* items;  // base of the admitted-items buffer
  int* lbuf;           // the second wave's ranks of the new items (kCmdRank)
  __device__ __forceinline__ void post(int lane, int cmd, int p1 = 0, int p2 = 0, int p3 = 0,
                                       int p4 = 0, int p5 = 0, int p6 = 0, int p7 = 0) {
    int4* c = reinterpret_cast<int4*>(ctl + 8 * (seq & 1));
    if (lane == 0) {
      c[0] = make_int4(cmd, p1, p2, p3);
      c[1] = make_int4(p4, p5, p6, p7);
    }
    ++seq;
    pair_sync();
  }
};

// One wave's share of sort_merge's rebuild walk when two waves share the
// query: rounds of 16 chunks from the top, the first wave (v = 0) rebuilding
// the upper 8 and the second the lower 8.  An output position reads an old
// position at most 256 (the new items of one pass) below it, i.e. at most 4
// chunks lower: inside this round's chunks of either wave or the top of the
// next round's.  So both waves finish a round's reads (barrier) before either
// writes, and a round's writes never meet reads still to come.
__device__ __forceinline__ void walk_pair(uint2* A, const uint2* X, const uint4* meta, int clo,
                                          int chi, int nout, int lane, int v) {
  const int nr = (chi - clo + 16) >> 4;
  for (int t = 0; t < nr; ++t) {
    const int c = chi - 16 * t - 8 * v;
    uint2 val[8];
    int pos[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cc = c - u;
      const bool okc = cc >= clo;
      const uint4 md = meta[(okc ? cc : clo) - clo];
      const int p = (cc << 6) + lane;
      const int bl = (int)md.z + (int)__builtin_amdgcn_mbcnt_hi(
                                     md.y, __builtin_amdgcn_mbcnt_lo(md.x, 0u));
      const bool mine = ((lane < 32 ? md.x : md.y) >> (lane & 31)) & 1u;
      const bool ok = okc && p < nout;
      const uint2* src = mine ? X + bl : A + (p - bl);
      val[u] = *(ok ? src : A);
      pos[u] = ok ? p : -1;
    }
    pair_sync();
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (pos[u] >= 0) A[pos[u]] = val[u];
  }
}

// The rebuild walk of sort_merge by one wave: output chunks chi down to lo
// (>= clo, the chunk of meta[0]), eight in flight.  Output chunk c reads only
// below 64 (c + 1) and writes [64 c, 64 (c + 1)), so going top-down the reads
// of lower chunks may run ahead of the writes, and the walk may stop at any
// chunk and resume below it later.
__device__ __forceinline__ void walk_all(uint2* A, const uint2* X, const uint4* meta, int clo,
                                         int chi, int nout, int lane, int lo = -1) {
  if (lo < clo) lo = clo;
  for (int c = chi; c >= lo; c -= 8) {
    uint4 md[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) md[u] = meta[max(c - u, lo) - clo];
    uint2 v[8];
    int pos[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cc = c - u;
      const int p = (cc << 6) + lane;
      // marks below this lane: mbcnt counts the set bits of the mask below it
      const int bl = (int)md[u].z + (int)__builtin_amdgcn_mbcnt_hi(
                                        md[u].y, __builtin_amdgcn_mbcnt_lo(md[u].x, 0u));
      const bool mine = ((lane < 32 ? md[u].x : md[u].y) >> (lane & 31)) & 1u;
      const bool ok = cc >= lo && p < nout;
      const uint2* src = mine ? X + bl : A + (p - bl);
      v[u] = *(ok ? src : A);
      pos[u] = ok ? p : -1;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (pos[u] >= 0) A[pos[u]] = v[u];
  }
}

// Item p of the array a merge makes while its rebuild walk has covered only
// the chunks above chi: the merge records of chunks clo..chi (see sort_merge)
// say whether position p takes a new item of X or an item of A moved up;
// positions below chunk clo keep A's, those above chi are rebuilt.  p is the
// same on every lane.
__device__ __forceinline__ uint2 merged_item(const uint2* A, const uint2* X, const uint4* meta,
                                             int clo, int chi, int p) {
  const int c = p >> 6;
  if (c < clo || c > chi) return A[p];
  const uint4 md = meta[c - clo];
  const int b = p & 63;
  const uint64_t mask = ((uint64_t)md.y << 32) | md.x;
  const int bl = (int)md.z + __popcll(mask & ((1ull << b) - 1ull));
  return ((mask >> b) & 1ull) ? X[bl] : A[p - bl];
}

// Ranks of the sorted new items X[0..m) (item r * 64 + lane in k[r], id[r])
// in A[0..n): L[r] = #{A items before it} by binary search (the used bit of A
// is ignored, so items being claimed meanwhile do not matter).  Then, unless
// no new item lands below nout, the merge records (see sort_merge) for the
// chunks from the smallest new item's position P0 up.  Returns P0.
template <int R>
__device__ __forceinline__ int rank_and_mark(const uint2* A, int n, int m, int nout,
                                             const float* k, const uint32_t* id, int* L,
                                             uint4* meta, int lane) {
#pragma unroll
  for (int r = 0; r < R; ++r) L[r] = 0;
  for (int step = 1 << (31 - __builtin_clz((unsigned)n)); step > 0; step >>= 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = L[r] + step - 1;
      if (r * 64 + lane < m && i < n) {
        const uint2 e = A[i];
        if (kid_less(__uint_as_float(e.x), e.y & ~kUsedBit, k[r], id[r])) L[r] += step;
      }
    }
  }
  const int P0 = __builtin_amdgcn_readlane(L[0], 0);
  if (P0 >= nout) return P0;
  const int clo = P0 >> 6, chi = (nout - 1) >> 6;
  for (int c = lane; c <= chi - clo; c += 64) meta[c] = make_uint4(0u, 0u, 0u, 0u);
  WSYNC();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * 64 + lane;
    const int p = L[r] + j;
    if (j < m && p < nout)
      atomicOr(reinterpret_cast<unsigned long long*>(&meta[(p >> 6) - clo]), 1ull << (p & 63));
  }
  WSYNC();
  const uint64_t below_lane = (1ull << lane) - 1;
  // new items below each chunk: an exclusive prefix over the chunks' counts,
  // 64 chunks at a time, bit-sliced through ballots
  int base = 0;
  for (int g = 0; g <= chi - clo; g += 64) {
    const int c = g + lane;
    const bool in = c <= chi - clo;
    const uint4 e = in ? meta[c] : make_uint4(0u, 0u, 0u, 0u);
    const int cnt = __popc(e.x) + __popc(e.y);
    int excl = base, tot = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) {
      const uint64_t mb = __ballot((cnt >> b) & 1);
      excl += __popcll(mb & below_lane) << b;
      tot += __popcll(mb) << b;
    }
    if (in) meta[c].z = (uint32_t)excl;
    base += tot;
  }
  WSYNC();
  return P0;
}

// Sorts the m (<= 64 R) items of X[0..m) by (key, id), writes them back in
// order and merges them into A[0..n), truncating at C.  Returns the position
// the smallest new item takes (>= C: none is kept).
//
// New item j lands at P_j = j + #{A items before it} (a binary search per
// item, all rows in step).  Per output chunk c of 64 positions, from the chunk
// of P_0 up, meta[c - clo] holds the bitmap of the new items' positions in
// the chunk (x, y) and the number of new items below it (z).  Output position
// p then takes new item #{new items below p} if marked, else A[p - #{new items
// below p}].  Output chunk c reads only below 64 (c + 1) and writes
// [64 c, 64 (c + 1)), so going top-down the reads of lower chunks may run ahead
// of the writes: eight chunks go in flight at once, with no lane exchange.
// after_sort() runs once X is sorted and A not yet changed.
//
// When after_sort() leaves nxt = {pa, kx}: the first unexpanded item of the
// merged array is either old A[pa] (pa < 0: none) or sorted new item kx (kx >=
// m: none), both unclaimed; its position goes to *first (nout when neither
// survives): A[pa] moves up by #{j : L_j <= pa}, item j lands at L_j + j.
//
// before_walk() runs once the records are built, before the rebuild (or in
// place of it when nothing lands).
//
// Two waves (PAIR) with `defer`: the rebuild is left to the caller, who gets
// (clo, chi, nout) of the walk in `pend`, or clo = -1 when nothing lands below
// nout.
template <int R, bool PAIR, class F, class G, class H>
__device__ __forceinline__ int sort_merge(uint2* A, int n_, int C_, uint2* X, int m_,
                                          uint4* meta, int lane, F&& after_sort, G&& lap,
                                          const int2& nxt, int* first, H&& before_walk,
                                          PairLink& link, bool defer, int3& pend) {
  const int n = __builtin_amdgcn_readfirstlane(n_);
  const int C = __builtin_amdgcn_readfirstlane(C_);
  const int m = __builtin_amdgcn_readfirstlane(m_);
  float k[R];
  uint32_t id[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * 64 + lane;
    if (j < m) {
      const uint2 v = X[j];
      k[r] = __uint_as_float(v.x);
      id[r] = v.y;
    } else {
      k[r] = kInf();
      id[r] = 0xffffffffu;
    }
  }
  bitonic_sort<R>(k, id, lane);
  WSYNC();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * 64 + lane;
    if (j < m) X[j] = make_uint2(__float_as_uint(k[r]), id[r]);
  }
  WSYNC();
  const int nout = min(C, n + m);
  if constexpr (PAIR) {
    // the second wave ranks the items and builds the merge records while
    // this one claims the next picks; the ranks come back in link.lbuf
    link.post(lane, kCmdRank, n, m, nout, (int)(X - link.items));
    after_sort();
    lap(4);
    link.post(lane, kCmdNop);
    const int* Lb = link.lbuf;
    if (first) {
      const int pa = __builtin_amdgcn_readfirstlane(nxt.x);
      const int kx = __builtin_amdgcn_readfirstlane(nxt.y);
      int fa = nout, fx = nout;
      if (pa >= 0) {
        int below = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = r * 64 + lane;
          below += __popcll(__ballot(j < m && Lb[j < m ? j : 0] <= pa));
        }
        fa = pa + below;
      }
      if (kx < m) fx = __builtin_amdgcn_readfirstlane(Lb[kx]) + kx;
      *first = min(nout, min(fa, fx));
    }
    const int P0 = __builtin_amdgcn_readfirstlane(Lb[0]);
    before_walk();
    lap(5);
    if (defer) pend = make_int3(P0 >= nout ? -1 : P0 >> 6, (nout - 1) >> 6, nout);
    if (P0 >= nout || defer) return P0;
    const int clo = P0 >> 6, chi = (nout - 1) >> 6;
    link.post(lane, kCmdWalk, clo, chi, nout, (int)(X - link.items));
    walk_pair(A, X, meta, clo, chi, nout, lane, 0);
    link.post(lane, kCmdNop);  // the second wave's chunks are written
    lap(6);
    return P0;
  }
  after_sort();
  lap(4);
  int L[R];
  const int P0 = rank_and_mark<R>(A, n, m, nout, k, id, L, meta, lane);
  if (first) {
    const int pa = __builtin_amdgcn_readfirstlane(nxt.x);
    const int kx = __builtin_amdgcn_readfirstlane(nxt.y);
    int fa = nout, fx = nout;
    if (pa >= 0) {
      int below = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) below += __popcll(__ballot(r * 64 + lane < m && L[r] <= pa));
      fa = pa + below;
    }
    if (kx < m) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((kx >> 6) == r) fx = __builtin_amdgcn_readlane(L[r], kx & 63) + kx;
    }
    *first = min(nout, min(fa, fx));
  }
  if (P0 >= nout) {
    before_walk();
    lap(5);
    return P0;
  }
  const int clo = P0 >> 6, chi = (nout - 1) >> 6;
  before_walk();
  lap(5);
  walk_all(A, X, meta, clo, chi, nout, lane);
  lap(6);
  return P0;
}

template <int D, int FUNC, int ST, int WIDE, bool PAIR>
__device__ __forceinline__ void beam_query(const SearchArgs& a, const WaveLds& w, const Marks& vis,
                           int qi, int k, int ef, int lane, int h, uint64_t ltm,
                           PairLink& link) {
  const DevIndex& ix = a.ix;
  const int C = max(ef, k);
  const int E = a.ent;  // entries per step: W * maxM0 <= E <= 512
  // distance path (PathNV): the wide kernel's, one vector per lane pair for
  // the first wave of a two-wave query, else the two-waves-per-SIMD one
  constexpr int DPATH = WIDE ? 1 : (PAIR ? 0 : 2);
  const int W = max(1, min(a.beam, E / ix.maxM0));
  const int erows = E >> 6;
  const uint32_t hslots = 2u * (uint32_t)E;
  // e / maxM0 = (e * mdiv) >> 16 exactly for e < 65536 / maxM0 (here e < 512,
  // maxM0 <= 64): the rounding error of mdiv adds less than e / 65536 < 1 / maxM0
  const int mdiv = (65536 + ix.maxM0 - 1) / ix.maxM0;
  uint2* A = w.A;
  unsigned long long nexp = 0, nd0 = 0, ndu = 1, nl = 0;

  int cur = ix.enter;
  float curd = dist_node<D, FUNC, ST>(w.qs, ix, cur, h);
  greedy_descent<D, FUNC, ST>(ix, w.qs, lane, h, cur, curd, ndu, nl);

  int n = 1;
  if (lane == 0) {
    A[0] = make_uint2(__float_as_uint(curd), (uint32_t)cur);
    vis.mark(cur);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  int ce = 0;
  // per-phase device time (100 MHz ticks) when counters are collected:
  // pick, adjacency + visited, distances, filter, sort + prediction, rank
  // search, rebuild, advance
  const bool timed = a.stats != nullptr;
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t tmark = timed ? __builtin_amdgcn_s_memrealtime() : 0;
  auto lap = [&](int p) __attribute__((always_inline)) {
    if (timed) {
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      ph[p] += now - tmark;
      tmark = now;
    }
  };
  // The next step's picks are the W smallest unexpanded items once this
  // step's admitted items are in: with capacity C = ef the survivors of the
  // truncation are a prefix of the array, so they are the W smallest of (the
  // first W unexpanded items of the array) and (the W smallest admitted
  // items), cut at the array's new last item.  As soon as the admitted items
  // are sorted, those W are claimed (marked expanded in place, in order in
  // pfq) and their adjacency rows (tn) load while the merge runs; their
  // visited marks (vn) load right after it.  The next step then takes the
  // surviving prefix of pfq without scanning the array.
  const bool claim = C == ef;
  // Two waves with a.defer: a merge's rebuild waits until the next step's
  // distances are handed to the second wave, and this wave runs it while they
  // are computed (the array is not read in between: the next step's picks
  // were claimed from the sorted items and the old array, and its last item
  // comes from the merge records).  The items of consecutive steps alternate
  // between two buffers, since the pending rebuild reads the last step's.
  const bool defer = PAIR && a.defer != 0;
  uint2* itb = w.items;
  int3 pend = make_int3(-1, 0, 0);  // pend.x >= 0: chunks pend.y down to pend.x of pX's walk are due
  const uint2* pX = w.items;
  auto flush = [&]() __attribute__((always_inline)) {
    if (pend.x >= 0) {
      walk_all(A, pX, w.meta, pend.x, pend.y, pend.z, lane);
      pend.x = -1;
      WSYNC();
    }
  };
  int pf_n = 0;  // > 0: pfq[0..pf_n) claimed for the next step
  int2 nxt = make_int2(-1, 0);  // first unclaimed items (see sort_merge)
  int first = 0;
  int tn[8], vn[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) tn[r] = vn[r] = -1;
  auto predict = [&](uint2* Xs, int nx) __attribute__((always_inline)) {
    int na = 0;  // up to W + 1: the first one left unclaimed is the next to expand
    for (int p = ce; p < n && na <= W; p += 128) {
      uint2 it[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int i = p + 64 * u + lane;
        it[u] = i < n ? A[i] : make_uint2(0u, kUsedBit);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const bool un = !(it[u].y & kUsedBit);
        const uint64_t um = __ballot(un);
        const int r = na + __popcll(um & ltm);
        if (un && r <= W) {
          w.pfa[r] = it[u];
          w.pfpos[r] = p + 64 * u + lane;
        }
        na = min(W + 1, na + __popcll(um));
      }
    }
    WSYNC();
    const int naw = min(W, na);
    const int nxw = min(W, nx);
    const int tot = naw + nxw;
    float ck = kInf();
    uint32_t ci = 0xffffffffu;
    if (lane < naw) {
      const uint2 v = w.pfa[lane];
      ck = __uint_as_float(v.x);
      ci = v.y;
    } else if (lane < tot) {
      const uint2 v = Xs[lane - naw];
      ck = __uint_as_float(v.x);
      ci = v.y;
    }
    int rank = 0;
    for (int j = 0; j < tot; ++j) {
      const float kj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ck), j));
      const uint32_t ij = (uint32_t)__builtin_amdgcn_readlane((int)ci, j);
      rank += kid_less(kj, ij, ck, ci);
    }
    const bool take = lane < tot && rank < W;
    if (take) {
      w.pfq[rank] = make_uint2(__float_as_uint(ck), ci);
      if (lane < naw) A[w.pfpos[lane]].y = ci | kUsedBit;
      else Xs[lane - naw].y = ci | kUsedBit;
    }
    pf_n = __builtin_amdgcn_readfirstlane(min(W, tot));
    // the claimed items of each side are a prefix of it
    const int ka = __popcll(__ballot(take && lane < naw));
    nxt = make_int2(ka < na ? w.pfpos[ka] : -1, pf_n - ka);
    WSYNC();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      tn[r] = -1;
      if (r < erows) {
        const int e = lane + 64 * r;
        const int li = (e * mdiv) >> 16;
        const int sl = e - li * ix.maxM0;
        if (li < pf_n)
          tn[r] = ix.l0[(int64_t)checked_node((int)w.pfq[li].y, (int)ix.n, a.err) * ix.l0stride + sl];
      }
    }
  };
  auto no_predict = [] {};
  // this step's marks land before any later read of them, then the marks of
  // the claimed picks' neighbours load (they arrive while the array is
  // rebuilt and the next picks are taken)
  auto load_marks = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
#pragma unroll
    for (int r = 0; r < 8; ++r) vn[r] = vis.raw(tn[r] >= 0 ? tn[r] : 0);
  };
  for (;;) {
    WSYNC();
    int nsel = 0;
    const bool claimed = pf_n > 0;
    const uint2 last = pend.x >= 0 ? merged_item(A, pX, w.meta, pend.x, pend.y, n - 1) : A[n - 1];
    if (claimed) {
      // the claimed picks that survived the last merge (a prefix of pfq)
      const uint2 q = lane < pf_n ? w.pfq[lane] : make_uint2(0u, 0u);
      const bool surv = lane < pf_n && !kid_less(__uint_as_float(last.x), last.y & ~kUsedBit,
                                                 __uint_as_float(q.x), q.y);
      nsel = __builtin_amdgcn_readfirstlane(__popcll(__ballot(surv)));
    } else {
      flush();
      // the first W unexpanded items below min(n, ef), in position order
      const int lim = min(n, ef);
      for (int p = ce; p < lim && nsel < W; p += 64) {
        const int i = p + lane;
        const bool un = i < lim && !(A[i].y & kUsedBit);
        const uint64_t um = __ballot(un);
        const int take = min(__popcll(um), W - nsel);
        const int r = __popcll(um & ltm);
        if (un && r < take) {
          const uint32_t y = A[i].y;
          A[i].y = y | kUsedBit;
          w.picks[nsel + r] = (int)y;
        }
        nsel += take;
      }
    }
    if (nsel == 0) break;
    nexp += nsel;
    const float topk = __uint_as_float(last.x);
    const int nsnap = n;
    // clear the dedupe table (2 E slots) while the picks land
    for (int i = lane; i < E; i += 64) reinterpret_cast<uint2*>(w.htab)[i] = make_uint2(~0u, ~0u);
    WSYNC();
    lap(0);

    // adjacency rows of the picks, entry e = 64 r + lane (W * maxM0 <= E),
    // and the visited marks of their entries: prefetched for claimed picks,
    // else loaded now, every row's loads issued before any is used (an id of
    // -1 reads mark 0 and is ignored)
    int t[8], vb[8];
    if (claimed) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int li = ((lane + 64 * r) * mdiv) >> 16;
        t[r] = li < nsel ? tn[r] : -1;
        vb[r] = vn[r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        t[r] = -1;
        if (r < erows) {
          const int e = lane + 64 * r;
          const int li = (e * mdiv) >> 16;
          const int sl = e - li * ix.maxM0;
          if (li < nsel)
            t[r] = ix.l0[(int64_t)checked_node(w.picks[li], (int)ix.n, a.err) * ix.l0stride + sl];
        }
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) vb[r] = vis.raw(t[r] >= 0 ? t[r] : 0);
    }
    bool fresh[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const bool valid = r < erows && t[r] >= 0;
      nl += __popcll(__ballot(valid));
      fresh[r] = valid && !vis.seen(vb[r], t[r]);
    }
    nl += nsel;
    // a neighbour shared by two picks is taken once (LDS hash, linear
    // probing).  Every row's compare-and-swap goes out in one round without
    // branching (a row with nothing to insert compares its slot against a
    // value no slot holds, which changes nothing); rows that met another id probe
    // the next slot in the next round.  Copies of one id follow the same probe
    // sequence, so exactly one of them inserts and the others see it.
    {
      uint32_t slot[8];
      bool pend[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        pend[r] = r < erows && fresh[r];
        slot[r] = (uint32_t)(((uint64_t)((uint32_t)t[r] * 2654435761u) * hslots) >> 32);
      }
      for (;;) {
        uint32_t old[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          old[r] = atomicCAS(&w.htab[slot[r]], pend[r] ? ~0u : 0xfffffffeu, (uint32_t)t[r]);
        bool more = false;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const bool ins = pend[r] && old[r] == ~0u;
          const bool dup = pend[r] && old[r] == (uint32_t)t[r];
          if (dup) fresh[r] = false;
          pend[r] = pend[r] && !ins && !dup;
          if (pend[r]) slot[r] = slot[r] + 1 == hslots ? 0u : slot[r] + 1;
          more |= pend[r];
        }
        if (__ballot(more) == 0) break;
      }
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (fresh[r]) vis.mark(t[r]);
    int ncand = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r < erows) {
        const uint64_t fm = __ballot(fresh[r]);
        if (fresh[r]) w.cand[ncand + __popcll(fm & ltm)] = t[r];
        ncand += __popcll(fm);
      }
    }
    nd0 += ncand;
    WSYNC();
    lap(1);
    // with two waves the second takes the candidates from h0 on, on the wide
    // path (several vectors in flight per lane pair; this wave keeps the
    // registers of the merge instead), both waves taking the same number of
    // rounds of loads; for FP32 vectors the second wave takes them all (this
    // wave's one-vector rounds are the slower ones there)
    int h0 = ncand;
    if constexpr (PAIR) {
      constexpr int P0 = ST == kF32 ? 0 : DistRound<D, FUNC, ST, DPATH>::value;
      constexpr int P1 = DistRound<D, FUNC, ST, 1>::value;
      const int r = (ncand + P0 + P1 - 1) / (P0 + P1);
      // deferred rebuild: the second wave takes every candidate (this one
      // rebuilds meanwhile; distances of its own, one vector per lane pair,
      // measured slower)
      h0 = defer ? 0 : max(0, ncand - P1 * r);
      if (h0 < ncand)
        link.post(lane, kCmdDist, ncand, h0, __float_as_int(topk), nsnap, ef, cur,
                  (int)(itb - w.items));
      if (defer && pend.x >= 0) {
        lap(2);
        flush();
        lap(6);
      }
    }
    dist_cands<D, FUNC, ST, DPATH>(w.qs, ix, w.cand, w.cdist, h0, lane, h, cur, a.err);
    WSYNC();
    lap(2);
    // keep what the snapshot admits; order does not matter (sorted next)
    int m = admit(w, 0, h0, topk, nsnap, ef, lane, ltm, itb);
    if constexpr (PAIR) {
      if (h0 < ncand) {
        link.post(lane, kCmdNop);  // the second wave's items are in
        const int m1 = __builtin_amdgcn_readfirstlane(w.ctl[16]);
        // they sit at items[h0..): move them down behind ours (m <= h0, so no
        // chunk overwrites an item a later chunk still reads)
        if (m < h0) {
          for (int b = 0; b < m1; b += 64) {
            const int i = b + lane;
            const uint2 v = i < m1 ? itb[h0 + i] : make_uint2(0u, 0u);
            if (i < m1) itb[m + i] = v;
          }
        }
        m += m1;
      }
    }
    WSYNC();
    lap(3);
    pf_n = 0;
    if (m > 0) {
      // sort by (distance, id) and merge, at most 256 items per pass (a later
      // pass only holds items the earlier ones do not); a single pass also
      // claims the next picks
      if (m <= 256 && claim) {
        auto pred = [&]() __attribute__((always_inline)) { predict(itb, m); };
        if (m <= 64) sort_merge<1, PAIR>(A, n, C, itb, m, w.meta, lane, pred, lap, nxt, &first,
                                         load_marks, link, defer, pend);
        else if (m <= 128) sort_merge<2, PAIR>(A, n, C, itb, m, w.meta, lane, pred, lap, nxt, &first,
                                               load_marks, link, defer, pend);
        else sort_merge<4, PAIR>(A, n, C, itb, m, w.meta, lane, pred, lap, nxt, &first,
                                 load_marks, link, defer, pend);
        n = min(C, n + m);
        ce = first;
        if (pend.x >= 0) {  // the rebuild reads these items: the next step takes the other buffer
          pX = itb;
          itb = itb == w.items ? w.items2 : w.items;
        }
      } else {
        for (int b = 0; b < m; b += 256) {
          const int cnt = min(256, m - b);
          int L0;
          if (cnt <= 64) L0 = sort_merge<1, PAIR>(A, n, C, itb + b, cnt, w.meta, lane, no_predict, lap, nxt,
                                          nullptr, no_predict, link, false, pend);
          else if (cnt <= 128) L0 = sort_merge<2, PAIR>(A, n, C, itb + b, cnt, w.meta, lane, no_predict, lap, nxt,
                                          nullptr, no_predict, link, false, pend);
          else L0 = sort_merge<4, PAIR>(A, n, C, itb + b, cnt, w.meta, lane, no_predict, lap, nxt,
                                          nullptr, no_predict, link, false, pend);
          n = min(C, n + cnt);
          if (L0 < ce) ce = L0;
          WSYNC();
        }
      }
    } else if (claim) {
      predict(itb, 0);
      load_marks();
      ce = nxt.x >= 0 ? nxt.x : n;
      lap(4);
    }
    WSYNC();
    if (pf_n == 0) {
      flush();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      ce = first_unused(A, ce, n, lane);
    }
    lap(7);
  }
  flush();
  WSYNC();
  if (timed && lane == 0)
    for (int p = 0; p < 8; ++p) atomicAdd(&a.stats->phase_ticks[p], ph[p]);
  const QueryCounters c{nexp, nd0, ndu, nl, 0};
  finish_query<D, FUNC, ST>(a, w, n, qi, k, lane, h, c);
}

// kCmdRank on the second wave: ranks of the sorted new items X[0..m) in
// A[0..n) into w.cand (the first wave reads them back) and the merge records.
// The first wave claims picks meanwhile, setting used bits in A and X, which
// neither the ranks nor the ids read here depend on.
template <int R>
__device__ __forceinline__ void helper_rank(const WaveLds& w, const uint2* X, int n, int m, int nout,
                                            int lane) {
  float k[R];
  uint32_t id[R];
  int L[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * 64 + lane;
    const uint2 v = j < m ? X[j] : make_uint2(0u, 0u);
    k[r] = __uint_as_float(v.x);
    id[r] = v.y & ~kUsedBit;
  }
  rank_and_mark<R>(w.A, n, m, nout, k, id, L, w.meta, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * 64 + lane;
    if (j < m) w.cand[j] = L[r];
  }
}

// The second wave of a two-wave query (kModeBeamPair): carries out the
// first wave's commands until it posts kCmdExit.
//   kCmdDist (ncand, h0, topk, nsnap, ef, fallback id, items buffer offset):
//     distances of cand[h0..ncand) and the admitted ones as items at
//     items[offset + h0..); their number goes to ctl[16];
//   kCmdRank (n, m, nout, offset of the new items in items): ranks and merge
//     records of a merge pass (helper_rank);
//   kCmdWalk (clo, chi, nout, offset of the new items in items): the lower
//     half of each round of the rebuild walk.
template <int D, int FUNC, int ST>
__device__ __forceinline__ void beam_helper(const SearchArgs& a, const WaveLds& w, int lane, int h,
                                            uint64_t ltm) {
  int seq = 0;
  for (;;) {
    pair_sync();
    const int4* c = reinterpret_cast<const int4*>(w.ctl + 8 * (seq & 1));
    ++seq;
    const int4 c0 = c[0], c1 = c[1];
    const int cmd = __builtin_amdgcn_readfirstlane(c0.x);
    if (cmd == kCmdExit) break;
    if (cmd == kCmdDist) {
      const int ncand = __builtin_amdgcn_readfirstlane(c0.y);
      const int h0 = __builtin_amdgcn_readfirstlane(c0.z);
      const float topk = __int_as_float(__builtin_amdgcn_readfirstlane(c0.w));
      const int nsnap = __builtin_amdgcn_readfirstlane(c1.x);
      const int ef = __builtin_amdgcn_readfirstlane(c1.y);
      const int fb = __builtin_amdgcn_readfirstlane(c1.z);
      uint2* itb = w.items + __builtin_amdgcn_readfirstlane(c1.w);
      dist_cands<D, FUNC, ST, 1>(w.qs, a.ix, w.cand + h0, w.cdist + h0, ncand - h0, lane, h, fb,
                                 a.err);
      WSYNC();
      const int m1 = admit(w, h0, ncand, topk, nsnap, ef, lane, ltm, itb + h0);
      if (lane == 0) w.ctl[16] = m1;
    } else if (cmd == kCmdRank) {
      const int n = __builtin_amdgcn_readfirstlane(c0.y);
      const int m = __builtin_amdgcn_readfirstlane(c0.z);
      const int nout = __builtin_amdgcn_readfirstlane(c0.w);
      const uint2* X = w.items + __builtin_amdgcn_readfirstlane(c1.x);
      if (m <= 64) helper_rank<1>(w, X, n, m, nout, lane);
      else if (m <= 128) helper_rank<2>(w, X, n, m, nout, lane);
      else helper_rank<4>(w, X, n, m, nout, lane);
    } else if (cmd == kCmdWalk) {
      walk_pair(w.A, w.items + __builtin_amdgcn_readfirstlane(c1.x), w.meta,
                __builtin_amdgcn_readfirstlane(c0.y), __builtin_amdgcn_readfirstlane(c0.z),
                __builtin_amdgcn_readfirstlane(c0.w), lane, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// Cooperative beam search (kModeBeamCoop; an extension with no reference
// counterpart): one query per block of kCoopWaves waves, each step split
// across them between workgroup barriers.  The step is beam_query's -- the
// first W unexpanded items of the array expanded together, the snapshot
// admission rule, the merge truncated at max(ef, k) -- so the results equal
// the beam model's:
//   1. picks and the first unexpanded item left (wave 0);
//   2. adjacency rows, visited marks and compaction, the entries split across
//      all threads: a mark is a bit, read from L2 first; where it is clear, an
//      atomic OR that returns the old word sets it, so the mark and the dedupe
//      across the picks' lists are one operation (two copies of an id: one
//      sees the bit clear);
//   3. distances (each wave a slice of the candidates, wide gathers) and the
//      snapshot admission, items placed through an LDS counter;
//   4. the rank of every new item among the new items (one item per thread:
//      keys are distinct (distance, id) pairs) and in the array (binary
//      search), the merge records, then the rebuild walk in rounds of 64
//      chunks, 8 per wave, every read of a round before any of its writes (a
//      position reads at most m positions below itself, so a round only reads
//      chunks that no earlier round has written).
// A whole CU serves one query, and beam widths up to 64 shorten the serial
// chain: this is the low-latency form.  Visited marks are always bits.
// The m new items of a cooperative step, items[0..m), sorted by (distance,
// id) into X: wave w sorts the run items[64 R w, 64 R (w + 1)) in registers
// (bitonic_sort) and writes it back; then an item's rank is its index in its
// run plus, for every other run, the number of that run's items before it
// (binary search; keys are distinct (distance, id) pairs).  Ends with a
// workgroup barrier.
template <int R>
__device__ __forceinline__ void coop_sort_new(uint2* items, uint2* X, int m, int wave, int lane) {
  constexpr int RUN = 64 * R;
  const int b0 = wave * RUN;
  float k[R];
  uint32_t id[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = b0 + r * 64 + lane;
    if (j < m) {
      const uint2 v = items[j];
      k[r] = __uint_as_float(v.x);
      id[r] = v.y;
    } else {
      k[r] = kInf();
      id[r] = 0xffffffffu;
    }
  }
  if (b0 < m) bitonic_sort<R>(k, id, lane);
  WSYNC();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = b0 + r * 64 + lane;
    if (j < m) items[j] = make_uint2(__float_as_uint(k[r]), id[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = b0 + r * 64 + lane;
    if (j < m) {
      int rank = r * 64 + lane;
      for (int b = 0; b < kCoopWaves; ++b) {
        const int s0 = b * RUN;
        if (b == wave || s0 >= m) continue;
        int lo = s0, hi = min(m, s0 + RUN);
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          const uint2 o = items[mid];
          if (kid_less(__uint_as_float(o.x), o.y, k[r], id[r])) lo = mid + 1; else hi = mid;
        }
        rank += lo - s0;
      }
      X[rank] = make_uint2(__float_as_uint(k[r]), id[r]);
    }
  }
  __syncthreads();
}

// The BF16 screen of the FP32 l2 search (an extension with no reference
// counterpart; cooperative kernel): a lower bound on the FP32 distance that
// vec_dist returns, from the BF16 copies q~, x~ of the query and the vector
// (both rounded to nearest even: |v - v~| <= 2^-8 |v| per coordinate), so a
// candidate the bound puts at or above the admission key is left out without
// its FP32 vector: the step admits exactly the items it admits without the
// screen.  With the FP32 sums sxq ~ q~.x~, sxx ~ |x~|^2 (D/4 dot2 steps per
// lane, then the pair's halves) and qq ~ |q~|^2, and u = 2^-24:
//   |q~ - x~|^2 >= (sxx - 2 sxq) + qq - E,  E = (D + 8) 2^-21 (|q~| + |x~|)^2
//     (the sums err by at most (D/2 + 8) u (|q~| + |x~|)^2 together);
//   |q - x| >= |q~ - x~| - |q - q~| - |x - x~| >= |q~ - x~| - eq - 2^-8 / (1 - 2^-8)
//     |x~|, with eq = |q - q~| measured once per query (make_query_bf16) and
//     the norms from sxx and qq enlarged by 2^-10 for their error;
//   vec_dist's FP32 sum of D non-negative rounded squares >= |q - x|^2
//     (1 - (D + 2) u).
// Each step is taken with slack for the FP32 rounding of the bound itself
// (the square root shrunk by 1e-5, the result by (D + 8) u + 2^-20).  A bound
// that is not a number never rejects.
template <int D>
__device__ __forceinline__ float screen_bound(const VecRegs<D, kBF16>& r, const float* qs, int h,
                                              int dstride) {
  constexpr int Q16 = D / 16;
  const uint32_t* qb = query_bf16(qs, dstride);
  float sxq = 0.f, sxx = 0.f;
#pragma unroll
  for (int c = 0; c < Q16; ++c) {
    const uint4 xr = r.c[c];
    const uint4 qr = *reinterpret_cast<const uint4*>(qb + 8 * c + 4 * h);
    sxq = dot2_bf16(xr.x, qr.x, sxq);
    sxq = dot2_bf16(xr.y, qr.y, sxq);
    sxq = dot2_bf16(xr.z, qr.z, sxq);
    sxq = dot2_bf16(xr.w, qr.w, sxq);
    sxx = dot2_bf16(xr.x, xr.x, sxx);
    sxx = dot2_bf16(xr.y, xr.y, sxx);
    sxx = dot2_bf16(xr.z, xr.z, sxx);
    sxx = dot2_bf16(xr.w, xr.w, sxx);
  }
#pragma unroll
  for (int t = 0; 2 * t + h < D % 16; ++t) {
    const int c = 16 * Q16 + 2 * t + h;
    const uint32_t xw = r.t[t];
    const uint32_t qw = (qb[c >> 1] >> (16 * (c & 1))) & 0xffffu;
    sxq = dot2_bf16(xw, qw, sxq);
    sxx = dot2_bf16(xw, xw, sxx);
  }
  sxq = sxq + __shfl_xor(sxq, 1);
  sxx = sxx + __shfl_xor(sxx, 1);
  const float qq = reinterpret_cast<const float*>(qb)[dstride / 2];
  const float eq = reinterpret_cast<const float*>(qb)[dstride / 2 + 1];
  const float nx = sqrtf(fmaxf(sxx, 0.f)) * (1.0f + 0x1p-10f);
  const float nsum = nx + sqrtf(fmaxf(qq, 0.f)) * (1.0f + 0x1p-10f);
  const float e = (float)(D + 8) * 0x1p-21f * nsum * nsum;
  const float dlo = ((sxx - 2.0f * sxq) + qq) - e;
  const float rr = dlo > 0.f ? sqrtf(dlo) * 0.99999f : 0.f;
  const float rp = rr - ((eq + nx * (0x1p-8f / (1.0f - 0x1p-8f))) * 1.00001f + 1e-30f);
  const float c3 = (float)(D + 8) * 0x1p-24f + 0x1p-20f;
  return rp > 0.f ? (rp * rp) * (1.0f - c3) : (rp == rp ? 0.f : rp);
}

// Screens the wave's candidates cand[s0..s1) against the admission key topk
// (the array full): the ones whose bound stays below topk are compacted, in
// order, to cand[s0..s0 + returned).  NV BF16 vectors per lane pair in
// flight.  verify: also computes each candidate's FP32 distance and sets bit
// 4 of the error word if a left-out one would have been admitted.
template <int D, int FUNC, int NV>
__device__ __forceinline__ int screen_slice(const SearchArgs& a, const WaveLds& w, int s0, int s1,
                                            float topk, int lane, int h, int fallback, bool verify) {
  const DevIndex& ix = a.ix;
  uint64_t ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int kept = 0;
  for (int b = s0; b < s1; b += 32 * NV) {
    VecRegs<D, kBF16> r[NV];
    int id[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int s = b + 32 * v + (lane >> 1);
      id[v] = checked_node(s < s1 ? w.cand[s] : fallback, (int)ix.n, a.err);
      vec_load<D, kBF16>(r[v], ix, id[v], h);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int s = b + 32 * v + (lane >> 1);
      const float lb = screen_bound<D>(r[v], w.qs, h, ix.dstride);
      const bool keep = !(lb >= topk);
      if (verify) {
        const float d = dist_node<D, FUNC, kF32>(w.qs, ix, id[v], h);
        if (s < s1 && !keep && d < topk && a.err) atomicOr(a.err, 4);
      }
      // lane 2j (h == 0) speaks for candidate s; the compaction keeps order
      const bool k0 = s < s1 && keep && h == 0;
      const uint64_t km = __ballot(k0);
      WSYNC();  // every read of this round's ids is done before the writes
      if (k0) w.cand[s0 + kept + __popcll(km & ltm)] = id[v];
      kept += __popcll(km);
    }
  }
  WSYNC();
  return kept;
}

template <int D, int FUNC, int ST, int DPATH>
__device__ __forceinline__ void coop_query(const SearchArgs& a, const WaveLds& w, const Marks& vis,
                                           int qi, int k, int ef, int wave, int lane, int h,
                                           uint64_t ltm) {
  constexpr int NT = kCoopWaves * kWave;
  constexpr int RMAX = 4;  // entries per thread and step: W * maxM0 <= 2048
  const DevIndex& ix = a.ix;
  const int tid = wave * kWave + lane;
  const int C = max(ef, k);
  const int W = max(1, min(a.beam, a.ent / ix.maxM0));
  uint2* A = w.A;
  int* ctl = w.ctl;
  int* picks = reinterpret_cast<int*>(w.pfa);  // [W <= 64] (pfa, pfq, pfpos: 104 words)
  int* wcnt = picks + 64;                      // [kCoopWaves] per-wave counts of a pick round
  // after the admission the candidates' region holds the step's new items,
  // sorted (X), and the admitted items' region, once sorted from, their ranks
  // in A (lrank)
  uint2* X = reinterpret_cast<uint2*>(w.cand);
  int* lrank = reinterpret_cast<int*>(w.items);
  uint32_t* vw = reinterpret_cast<uint32_t*>(vis.p);
  unsigned long long ndu = 1, nl = 0;  // wave 0: the descent's counters and rows read
  uint32_t nexp = 0, nd0 = 0;           // expansions and level-0 distances of the query
  // the BF16 screen: FP32 l2 on a specialised dimension
  constexpr bool kScreen = ST == kF32 && (FUNC == kL2Sqr16 || FUNC == kL2Sqr) && D > 0 && D % 4 == 0;
  uint32_t nscr = 0, npass = 0;  // this wave's screened candidates, and passed

  if (wave == 0) {
    int cur = ix.enter;
    float curd = dist_node<D, FUNC, ST>(w.qs, ix, cur, h);
    greedy_descent<D, FUNC, ST>(ix, w.qs, lane, h, cur, curd, ndu, nl);
    if (lane == 0) {
      A[0] = make_uint2(__float_as_uint(curd), (uint32_t)cur);
      vis.mark(cur);
      ctl[4] = cur;
      ctl[6] = 0;
    }
  }
  __syncthreads();
  const int fallback = ctl[4];
  int n = 1, ce = 0;
  // per-phase device time (100 MHz ticks, taken by thread 0 into ctl[10..18)
  // when counters are collected): picks, rows + marks, distances +
  // admission, ranks among the new items, ranks in the array, merge records,
  // rebuild, advance
  const bool timed = a.stats != nullptr && tid == 0;
  uint64_t tmark = 0;
  if (timed) {
    for (int p = 0; p < 8; ++p) ctl[10 + p] = 0;
    tmark = __builtin_amdgcn_s_memrealtime();
  }
  auto lap = [&](int p) __attribute__((always_inline)) {
    if (timed) {
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      ctl[10 + p] += (int)(now - tmark);
      tmark = now;
    }
  };
  for (;;) {
    // 1. the first W unexpanded items below min(n, ef), in position order,
    // marked expanded, and the first unexpanded item left (pa, n if none):
    // rounds of 512 positions from ce, one chunk of 64 per wave, the order of
    // an unexpanded item among those found taken from the waves' counts
    const int lim = min(n, ef);
    if (tid == 0) {
      ctl[2] = 0;  // picks
      ctl[7] = n;  // pa
      ctl[0] = 0;  // candidates
      ctl[1] = 0;  // admitted items
    }
    int found = 0;  // unexpanded items in the rounds before this one
    for (int base = ce; base < n; base += NT) {
      const int i = base + tid;
      const bool un = i < n && !(A[i].y & kUsedBit);
      const uint64_t um = __ballot(un);
      if (lane == 0) wcnt[wave] = __popcll(um);
      __syncthreads();
      int before = 0, total = 0;
#pragma unroll
      for (int v = 0; v < kCoopWaves; ++v) {
        const int c = wcnt[v];
        before += v < wave ? c : 0;
        total += c;
      }
      const int r = found + before + __popcll(um & ltm);
      const bool pick = un && i < lim && r < W;
      if (pick) {
        const uint32_t y = A[i].y;
        A[i].y = y | kUsedBit;
        picks[r] = (int)y;
      }
      // one LDS atomic per wave: its picks, and its first unexpanded item left
      const uint64_t pm = __ballot(pick), lm = __ballot(un && !pick);
      if (lane == 0) {
        if (pm) atomicAdd(&ctl[2], __popcll(pm));
        if (lm) atomicMin(&ctl[7], base + wave * kWave + __ffsll((long long)lm) - 1);
      }
      found += total;
      __syncthreads();
      if (ctl[7] < n) break;  // every later item is past lim or past the W picks
    }
    if (wave == 0) nl += ctl[2];
    __syncthreads();
    lap(0);
    const int nsel = ctl[2];
    if (nsel == 0) break;
    nexp += nsel;
    const float topk = akey(A, n - 1);
    const int nsnap = n;
    const int pa = ctl[7];

    // 2. rows, marks and fresh entries: entry e = tid + NT r of the nsel
    // lists of maxM0 slots; every row load, then every mark, in flight at once
    const int nent = nsel * ix.maxM0;
    const int R = (nent + NT - 1) / NT;
    int t[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int e = tid + NT * r;
      t[r] = -1;
      if (r < R && e < nent) {
        const int li = e / ix.maxM0;
        const int sl = e - li * ix.maxM0;
        t[r] = ix.l0[(int64_t)checked_node(picks[li], (int)ix.n, a.err) * ix.l0stride + sl];
      }
    }
    // the mark words read from L2 first: about half the entries are visited
    // already, and only the others take the atomic OR (whose old word decides,
    // so two copies of an id in this step still yield one fresh entry)
    uint32_t pre[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r)
      pre[r] = t[r] >= 0 ? __hip_atomic_load(vw + (t[r] >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : 0xffffffffu;
    uint32_t old[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      old[r] = 0xffffffffu;
      if (t[r] >= 0 && !((pre[r] >> (t[r] & 31)) & 1u))
        old[r] = atomicOr(vw + (t[r] >> 5), 1u << (t[r] & 31));
    }
    int lnl = 0;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      if (r < R) {
        lnl += __popcll(__ballot(t[r] >= 0));
        const bool fresh = t[r] >= 0 && !((old[r] >> (t[r] & 31)) & 1u);
        const uint64_t fm = __ballot(fresh);
        int base = 0;
        if (lane == 0 && fm) base = atomicAdd(&ctl[0], __popcll(fm));
        base = __shfl(base, 0);
        if (fresh) w.cand[base + __popcll(fm & ltm)] = t[r];
      }
    }
    if (lane == 0 && lnl) atomicAdd(&ctl[6], lnl);
    __syncthreads();
    lap(1);
    const int ncand = ctl[0];
    nd0 += ncand;

    // 3. distances of a slice of the candidates per wave, and the admission
    // (with the array full, the BF16 screen first leaves out the candidates
    // its bound proves above topk)
    const int per = (ncand + kCoopWaves - 1) / kCoopWaves;
    const int s0 = min(ncand, wave * per), s1 = min(ncand, s0 + per);
    int se = s1;
    if constexpr (kScreen) {
      if (a.screen && nsnap >= ef) {
        // one BF16 vector per lane pair in flight at four waves per SIMD (two
        // spill), two at two waves per SIMD
        se = s0 + screen_slice<D, FUNC, DPATH == 1 ? 2 : 1>(a, w, s0, s1, topk, lane, h, fallback,
                                                             a.screen == 2);
        nscr += s1 - s0;
        npass += se - s0;
      }
    }
    dist_cands<D, FUNC, ST, DPATH>(w.qs, ix, w.cand + s0, w.cdist + s0, se - s0, lane, h, fallback,
                                   a.err);
    WSYNC();
    for (int b = s0; b < se; b += 64) {
      const int s = b + lane;
      const float dj = s < se ? w.cdist[s] : 0.f;
      const bool keep = s < se && (dj < topk || nsnap < ef);
      const uint64_t km = __ballot(keep);
      int base = 0;
      if (lane == 0 && km) base = atomicAdd(&ctl[1], __popcll(km));
      base = __shfl(base, 0);
      if (keep)
        w.items[base + __popcll(km & ltm)] = make_uint2(__float_as_uint(dj), (uint32_t)w.cand[s]);
    }
    __syncthreads();
    lap(2);
    const int m = ctl[1];
    if (m == 0) {
      ce = pa;
      continue;
    }

    // 4. ranks among the new items (sorted into X), then in the array
    if (m <= 512) coop_sort_new<1>(w.items, X, m, wave, lane);
    else if (m <= 1024) coop_sort_new<2>(w.items, X, m, wave, lane);
    else coop_sort_new<4>(w.items, X, m, wave, lane);
    lap(3);
    for (int j = tid; j < m; j += NT) {
      const uint2 v = X[j];
      lrank[j] = lower_count_kid(A, n, __uint_as_float(v.x), v.y);
    }
    __syncthreads();
    lap(4);
    const int nout = min(C, n + m);
    const int P0 = lrank[0];
    if (P0 < nout) {
      const int clo = P0 >> 6, chi = (nout - 1) >> 6;
      for (int c = tid; c <= chi - clo; c += NT) w.meta[c] = make_uint4(0u, 0u, 0u, 0u);
      __syncthreads();
      for (int j = tid; j < m; j += NT) {
        const int p = lrank[j] + j;
        if (p < nout)
          atomicOr(reinterpret_cast<unsigned long long*>(&w.meta[(p >> 6) - clo]), 1ull << (p & 63));
      }
      __syncthreads();
      if (wave == 0) {
        // new items below each chunk: an exclusive prefix over the chunks'
        // counts, 64 chunks at a time, bit-sliced through ballots
        const uint64_t below_lane = (1ull << lane) - 1;
        int base = 0;
        for (int g = 0; g <= chi - clo; g += 64) {
          const int c = g + lane;
          const bool in = c <= chi - clo;
          const uint4 e = in ? w.meta[c] : make_uint4(0u, 0u, 0u, 0u);
          const int cnt = __popc(e.x) + __popc(e.y);
          int excl = base, tot = 0;
#pragma unroll
          for (int b = 0; b < 7; ++b) {
            const uint64_t mb = __ballot((cnt >> b) & 1);
            excl += __popcll(mb & below_lane) << b;
            tot += __popcll(mb) << b;
          }
          if (in) w.meta[c].z = (uint32_t)excl;
          base += tot;
        }
      }
      __syncthreads();
      lap(5);
      // the rebuild: rounds of 64 chunks from the top, 8 per wave
      for (int c0 = chi; c0 >= clo; c0 -= 64) {
        uint2 val[8];
        int pos[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int cc = c0 - 8 * wave - u;
          const bool okc = cc >= clo;
          const uint4 md = w.meta[(okc ? cc : clo) - clo];
          const int p = (cc << 6) + lane;
          const int bl = (int)md.z + (int)__builtin_amdgcn_mbcnt_hi(
                                         md.y, __builtin_amdgcn_mbcnt_lo(md.x, 0u));
          const bool mine = ((lane < 32 ? md.x : md.y) >> (lane & 31)) & 1u;
          const bool ok = okc && p < nout;
          const uint2* src = mine ? X + bl : A + (p - bl);
          val[u] = *(ok ? src : A);
          pos[u] = ok ? p : -1;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (pos[u] >= 0) A[pos[u]] = val[u];
        __syncthreads();
      }
    }
    lap(6);
    // the first unexpanded item of the merged array: the old one left at pa
    // (moved up by the new items before it) or the smallest new item
    if (tid == 0) {
      int pan = nout;
      if (pa < n) {
        int lo = 0, hi = m;  // new items with a rank <= pa (ranks are nondecreasing)
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (lrank[mid] <= pa) lo = mid + 1; else hi = mid;
        }
        pan = pa + lo;
      }
      ctl[9] = min(nout, min(pan, P0));
    }
    __syncthreads();
    lap(7);
    n = nout;
    ce = ctl[9];
  }
  if (timed)
    for (int p = 0; p < 8; ++p) atomicAdd(&a.stats->phase_ticks[p], (unsigned long long)ctl[10 + p]);
  if (kScreen && a.stats && lane == 0 && nscr) {
    atomicAdd(&a.stats->screened, (unsigned long long)nscr);
    atomicAdd(&a.stats->screen_pass, (unsigned long long)npass);
  }
  if (wave == 0) {
    const QueryCounters c{nexp, nd0, ndu, nl + (unsigned long long)ctl[6], 0};
    finish_query<D, FUNC, ST>(a, w, n, qi, k, lane, h, c);
  }
}

// One query per block of kCoopWaves waves (coop_query).
template <int D, int FUNC, int ST, int DPATH>
__device__ __forceinline__ void coop_body(const SearchArgs& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NT = kCoopWaves * kWave;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane & 1;
  const uint64_t ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const WaveLds w = carve(smem, a.cap, a.ix.dstride, a.ent);
  const int64_t slot = a.slot_base + (int64_t)blockIdx.x;
  uint8_t* vslot = a.ix.visited + slot * a.ix.vstride;
  const Marks vis{vslot, 0, true};
  const int dim = a.ix.dim;
  const int64_t nq = a.nq_dev ? *a.nq_dev : a.nq;
  for (;;) {
    if (threadIdx.x == 0) w.ctl[3] = atomicAdd(a.work, 1);
    __syncthreads();
    int qi = w.ctl[3];
    if (qi >= nq) break;
    if (a.order) qi = a.order[qi];
    const int k = a.k_of ? a.k_of[qi] : a.k;
    const int ef = a.ef_of ? a.ef_of[qi] : a.ef;
    const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
    // clear the slot's bit marks (the last query's ORs have completed); the
    // query into LDS, normalised in place for cosine (hnsw.h:486-497)
    uint4* v4 = reinterpret_cast<uint4*>(vslot);
    const int64_t nv4 = (a.ix.n + 128) / 128;
    for (int64_t i = threadIdx.x; i < nv4; i += NT) v4[i] = make_uint4(0, 0, 0, 0);
    const float* qg = a.queries + (int64_t)qi * dim;
    for (int i = threadIdx.x; i < a.ix.dstride; i += NT) w.qs[i] = i < dim ? qg[i] : 0.f;
    wait_vmem();      // the clearing stores are performed in L2 before any wave's OR
    __syncthreads();
    if (wave == 0) {
      if constexpr (FUNC == kNormCos) {
        float inv = 0.f;
        if (lane == 0) {
          float sum = 0.f;
          for (int i = 0; i < dim; ++i) sum = sum + w.qs[i] * w.qs[i];
          inv = sum != 0.0f ? 1.0f / sqrtf(sum) : 0.f;
        }
        inv = __shfl(inv, 0);
        if (inv != 0.f)
          for (int i = lane; i < dim; i += 64) w.qs[i] = w.qs[i] * inv;
        WSYNC();
      }
      if (ST == kBF16 || a.screen) make_query_bf16(w.qs, a.ix.dstride, lane);
    }
    __syncthreads();
    if (k <= 0 || a.ix.n == 0) {
      if (wave == 0) {
        for (int p = lane; p < k; p += 64) {
          a.out_ids[(int64_t)qi * a.out_stride + p] = -1;
          a.out_dists[(int64_t)qi * a.out_stride + p] = 0.f;
        }
        if (lane == 0) a.out_counts[qi] = 0;
      }
      __syncthreads();
      continue;
    }
    coop_query<D, FUNC, ST, DPATH>(a, w, vis, qi, k, ef, wave, lane, h, ltm);
    __syncthreads();
    if (a.q_ticks && threadIdx.x == 0)
      a.q_ticks[qi] = (uint32_t)(__builtin_amdgcn_s_memrealtime() - t_start);
  }
}

// ---------------------------------------------------------------------------
// SearchOld at level 0 (hnsw_distfunc_opt.cc:99-148): a candidate heap
// ordered by -distance (unbounded, in this slot's HBM scratch), a result heap
// of at most ef items and the k-item result queue.  Admission of each fresh
// neighbour depends on the heap state the previous one left, so the wave
// computes the distances of an expansion in parallel and lane 0 replays the
// admissions in list order with std::priority_queue's heap steps.  The LDS
// sorted-array space holds the ef-heap in [0, ef] and the result queue
// (sorted by (distance, internal id), the order KNNQueue pops in) after it.
template <int D, int FUNC, int ST>
__device__ __forceinline__ void old_query(const SearchArgs& a, const WaveLds& w, const Marks& vis, int qi,
                          int k, int ef, int lane, int h, uint64_t ltm, int64_t slot) {
  const DevIndex& ix = a.ix;
  unsigned long long nexp = 0, nd0 = 0, ndu = 1, nl = 0;
  float* hk = a.old_heap_key + slot * a.old_heap_cap;   // candidate heap keys (-d)
  int32_t* hi = a.old_heap_id + slot * a.old_heap_cap;
  // the wave's sorted-array storage (cap >= ef + 1 + k items of 8 bytes)
  float* bk = reinterpret_cast<float*>(w.A);            // ef-heap (max on d)
  int32_t* bi = reinterpret_cast<int32_t*>(bk + (ef + 1));
  float* rk = reinterpret_cast<float*>(bi + (ef + 1));  // result queue, ascending
  int32_t* ri = reinterpret_cast<int32_t*>(rk + k);

  int cur = ix.enter;
  float curd = dist_node<D, FUNC, ST>(w.qs, ix, cur, h);
  greedy_descent<D, FUNC, ST>(ix, w.qs, lane, h, cur, curd, ndu, nl);

  // lane 0 owns the heap sizes; they are broadcast when the wave needs them
  int hs = 0, bs = 0, rs = 0;
  bool overflow = false;
  auto offer_result = [&](float d, int id) {  // KNNQuery::CheckAndAddToResult
    if (rs < k || d < rk[rs - 1]) {
      int p = rs < k ? rs : k - 1;  // slot freed by the popped maximum
      if (rs < k) ++rs;
      while (p > 0 && (rk[p - 1] > d || (rk[p - 1] == d && ri[p - 1] > id))) {
        rk[p] = rk[p - 1];
        ri[p] = ri[p - 1];
        --p;
      }
      rk[p] = d;
      ri[p] = id;
    }
  };
  if (lane == 0) {
    hk[0] = -curd;
    hi[0] = cur;
    hs = 1;
    bk[0] = curd;
    bi[0] = cur;
    bs = 1;
    offer_result(curd, cur);
    vis.mark(cur);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  for (;;) {
    int node = -1;
    if (lane == 0) {
      if (hs > 0 && !(-hk[0] > bk[0])) {
        node = hi[0];
        emu_heap_pop(hk, hi, hs);
        --hs;
      }
    }
    node = __shfl(node, 0);
    if (node < 0) break;
    ++nexp;
    const int32_t* lst = ix.l0 + (int64_t)node * ix.l0stride;
    const int t = lane < ix.maxM0 ? lst[lane] : -1;
    const bool valid = t >= 0;
    nl += 1 + __popcll(__ballot(valid));
    bool fresh = false;
    if (valid) fresh = !vis.seen(t);
    if (ix.has_dups) {
      w.cand[lane] = t;
      WSYNC();
      for (int j = 0; j < lane; ++j)
        if (w.cand[j] == t) fresh = false;
      WSYNC();
    }
    if (fresh) vis.mark(t);
    const uint64_t fm = __ballot(fresh);
    const int ncand = __popcll(fm);
    if (fresh) w.cand[__popcll(fm & ltm)] = t;
    nd0 += ncand;
    WSYNC();
    dist_cands<D, FUNC, ST, 3>(w.qs, ix, w.cand, w.cdist, ncand, lane, h, node, a.err);
    WSYNC();
    if (lane == 0) {
      for (int j = 0; j < ncand; ++j) {
        const float d = w.cdist[j];
        const int id = w.cand[j];
        if (bk[0] > d || bs < ef) {
          if (hs >= a.old_heap_cap) {
            overflow = true;
            break;
          }
          hk[hs] = -d;
          hi[hs] = id;
          ++hs;
          emu_heap_push(hk, hi, hs);
          offer_result(d, id);
          bk[bs] = d;
          bi[bs] = id;
          ++bs;
          emu_heap_push(bk, bi, bs);
          if (bs > ef) {
            emu_heap_pop(bk, bi, bs);
            --bs;
          }
        }
      }
    }
    if (__shfl((int)overflow, 0)) break;
    WSYNC();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  }
  WSYNC();
  rs = __shfl(rs, 0);
  if (__shfl((int)overflow, 0) && lane == 0 && a.err) atomicOr(a.err, 1);
  const int64_t row = (int64_t)qi * a.out_stride;
  for (int p = lane; p < k; p += 64) {
    if (p < rs) {
      a.out_ids[row + p] = ix.ext_id[ri[p]];
      a.out_dists[row + p] = rk[p];
      if (a.out_internal) a.out_internal[row + p] = ri[p];
    } else {
      a.out_ids[row + p] = -1;
      a.out_dists[row + p] = 0.f;
      if (a.out_internal) a.out_internal[row + p] = -1;
    }
  }
  if (lane == 0) {
    a.out_counts[qi] = rs;
    if (a.stats) {
      atomicAdd(&a.stats->expansions, nexp);
      atomicAdd(&a.stats->dist_level0, nd0);
      atomicAdd(&a.stats->dist_upper, ndu);
      atomicAdd(&a.stats->list_ints_read, nl);
    }
  }
}

// MODE = kModeAll runs the reference's searches (V1Merge, SearchOld by
// algoType / ef).  kModeBeam / kModeBeamWide run the beam search for every
// query of the batch, each kernel carrying only its own loop's register
// state; kModeBeamWide keeps a round's vector loads in flight together (one
// wave per SIMD).  kModeBeamPair gives each query a block of two waves: the
// first runs the search, the second computes each step's distances (with the
// deferred rebuild) or their later rounds and half of the rebuild
// (beam_helper).
template <int D, int FUNC, int ST, int MODE>
__device__ __forceinline__ void search_body(const SearchArgs& a, int lds_wave_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane & 1;
  const uint64_t ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // two-wave queries: one query per block, both waves on the same LDS
  constexpr bool PAIR = MODE == kModeBeamPair;
  const WaveLds w = carve(smem + (PAIR ? 0 : wave * lds_wave_bytes), a.cap, a.ix.dstride, a.ent);
  if constexpr (PAIR) {
    if (wave == 1) {
      beam_helper<D, FUNC, ST>(a, w, lane, h, ltm);
      return;
    }
  }
  PairLink link{w.ctl, 0, w.items, w.cand};
  const int64_t slot = a.slot_base + (PAIR ? (int64_t)blockIdx.x
                                           : (int64_t)blockIdx.x * (blockDim.x >> 6) + wave);
  uint8_t* vslot = a.ix.visited + slot * a.ix.vstride;
  const int dim = a.ix.dim;
  const int64_t nq = a.nq_dev ? *a.nq_dev : a.nq;

  for (;;) {
    int qi = 0;
    if (lane == 0) qi = atomicAdd(a.work, 1);
    qi = __shfl(qi, 0);
    if (qi >= nq) break;
    if (a.order) qi = a.order[qi];  // uneven batches: largest ef first
    const int k = a.k_of ? a.k_of[qi] : a.k;
    const int ef = a.ef_of ? a.ef_of[qi] : a.ef;
    const uint64_t t_start = __builtin_amdgcn_s_memrealtime();

    // a fresh epoch for this slot's visited marks (hnsw.h:576-583), or bit
    // marks cleared once the last query's ORs have completed (beam mode only:
    // the exact searches keep epoch bytes, their code free of the choice)
    Marks vis{vslot, 0, MODE != kModeAll && a.vbits != 0};
    if (vis.bits) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      uint4* v4 = reinterpret_cast<uint4*>(vslot);
      const int64_t nv4 = (a.ix.n + 128) / 128;  // 32 marks per word, 4 words per store
      for (int64_t i = lane; i < nv4; i += 64) v4[i] = make_uint4(0, 0, 0, 0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    } else {
      int epi = 0;
      if (lane == 0) epi = a.ix.epoch[slot];
      epi = __shfl(epi, 0);
      uint8_t ep = (uint8_t)(epi + 1);
      if (ep == 0) {
        uint4* v4 = reinterpret_cast<uint4*>(vslot);
        const int64_t nv4 = a.ix.vstride / 16;
        for (int64_t i = lane; i < nv4; i += 64) v4[i] = make_uint4(0, 0, 0, 0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        ep = 1;
      }
      if (lane == 0) a.ix.epoch[slot] = ep;
      vis.ep = ep;
    }

    // the query, normalised in place for cosine (hnsw.h:486-497)
    const float* qg = a.queries + (int64_t)qi * dim;
    for (int i = lane; i < a.ix.dstride; i += 64) w.qs[i] = i < dim ? qg[i] : 0.f;
    WSYNC();
    if constexpr (FUNC == kNormCos) {
      float inv = 0.f;
      if (lane == 0) {
        float sum = 0.f;
        for (int i = 0; i < dim; ++i) sum = sum + w.qs[i] * w.qs[i];
        inv = sum != 0.0f ? 1.0f / sqrtf(sum) : 0.f;
      }
      inv = __shfl(inv, 0);
      if (inv != 0.f)
        for (int i = lane; i < dim; i += 64) w.qs[i] = w.qs[i] * inv;
      WSYNC();
    }
    if constexpr (ST == kBF16) {
      make_query_bf16(w.qs, a.ix.dstride, lane);
      WSYNC();
    }
    if (k <= 0 || a.ix.n == 0) {
      for (int p = lane; p < k; p += 64) {
        a.out_ids[(int64_t)qi * a.out_stride + p] = -1;
        a.out_dists[(int64_t)qi * a.out_stride + p] = 0.f;
      }
      if (lane == 0) a.out_counts[qi] = 0;
      continue;
    }
    // Hnsw::Search dispatch (hnsw.cc:724): SearchOld for algoType=old, or
    // hybrid with ef >= 1000
    if constexpr (MODE == kModeAll) {
      const bool use_old = a.algo == kAlgoOld || (a.algo == kAlgoHybrid && ef >= 1000);
      if (use_old)
        old_query<D, FUNC, ST>(a, w, vis, qi, k, ef, lane, h, ltm, slot);
      else
        v1merge_query<D, FUNC, ST>(a, w, vis, qi, k, ef, lane, h, ltm);
    } else {
      beam_query<D, FUNC, ST, MODE == kModeBeamWide, PAIR>(a, w, vis, qi, k, ef, lane, h, ltm,
                                                          link);
    }
    WSYNC();
    if (a.q_ticks && lane == 0)
      a.q_ticks[qi] = (uint32_t)(__builtin_amdgcn_s_memrealtime() - t_start);
  }
  if constexpr (PAIR) link.post(lane, kCmdExit);
}

// kModeAll / kModeBeam: register budget of two waves per SIMD
template <int D, int FUNC, int ST, int MODE>
__global__ void __launch_bounds__(kWave* kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(2)))
    hnsw_search_kernel(SearchArgs a, int lds_wave_bytes) {
  search_body<D, FUNC, ST, MODE>(a, lds_wave_bytes);
}

// kModeBeamWide: the whole register file of one wave per SIMD
template <int D, int FUNC, int ST>
__global__ void __launch_bounds__(kWave* kWavesPerBlock)
    hnsw_search_kernel_wide(SearchArgs a, int lds_wave_bytes) {
  search_body<D, FUNC, ST, kModeBeamWide>(a, lds_wave_bytes);
}

// kModeBeamPair: blocks of two waves sharing one query, two waves per SIMD
template <int D, int FUNC, int ST>
__global__ void __launch_bounds__(2 * kWave) __attribute__((amdgpu_waves_per_eu(2)))
    hnsw_search_kernel_pair(SearchArgs a, int lds_wave_bytes) {
  search_body<D, FUNC, ST, kModeBeamPair>(a, lds_wave_bytes);
}

// kModeBeamCoop: one query per block of kCoopWaves waves, two waves per SIMD
// (one block per CU), the wide distance path
template <int D, int FUNC, int ST>
__global__ void __launch_bounds__(kCoopWaves * kWave) __attribute__((amdgpu_waves_per_eu(2)))
    hnsw_search_kernel_coop(SearchArgs a, int) {
  coop_body<D, FUNC, ST, 1>(a);
}

// kModeBeamCoop2: the same with the register file of four waves per SIMD, so
// two blocks share a CU where LDS holds two queries; one FP32 vector (two BF16,
// four FP8) per lane pair in flight
template <int D, int FUNC, int ST>
__global__ void __launch_bounds__(kCoopWaves * kWave) __attribute__((amdgpu_waves_per_eu(4)))
    hnsw_search_kernel_coop2(SearchArgs a, int) {
  coop_body<D, FUNC, ST, ST == kF32 ? 0 : 2>(a);
}

using KernelFn = void (*)(SearchArgs, int);

template <int D, int FUNC, int ST, int MODE>
KernelFn kernel_of() {
  if constexpr (MODE == kModeBeamWide) return hnsw_search_kernel_wide<D, FUNC, ST>;
  else if constexpr (MODE == kModeBeamPair) return hnsw_search_kernel_pair<D, FUNC, ST>;
  else if constexpr (MODE == kModeBeamCoop) return hnsw_search_kernel_coop<D, FUNC, ST>;
  else if constexpr (MODE == kModeBeamCoop2) return hnsw_search_kernel_coop2<D, FUNC, ST>;
  else return hnsw_search_kernel<D, FUNC, ST, MODE>;
}

}  // namespace nmsamd


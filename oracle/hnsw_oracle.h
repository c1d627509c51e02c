/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement of nmslib's optimized-HNSW query path, written to be the
 * checker of the HIP product path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Each function follows a reference routine (files relative to
 * similarity_search/ of nmslib 2.1.2):
 *   ora_load                src/method/hnsw.cc:993-1074 (LoadIndex/LoadOptimizedIndex),
 *                           layout include/method/hnsw.h:328-371, include/object.h:40-104
 *   ora_search              src/method/hnsw_distfunc_opt.cc:46-283 (SearchOld, SearchV1Merge),
 *                           dispatch src/method/hnsw.cc:717-746,
 *                           result heap include/knnqueue.h:55-74, src/knnquery.cc:67-75
 *   sorted array            include/sort_arr_bi.h:61-199
 *   ora_dist                include/method/hnsw_distfunc_opt_impl_inline.h:42-173 (AVX order),
 *                           L1NormSIMD / LInfNormSIMD src/distcomp_lp.cc:78-247 (SSE order),
 *                           src/method/hnsw.cc:70-102
 *   ora_normalize           include/method/hnsw.h:486-497
 *   ora_brute_force         src/method/seqsearch.cc:143-177 ranked as
 *                           include/gold_standard.h:73-76 (distance, then id), in FP64
 */
#ifndef NMSLIB_AMD_HNSW_ORACLE_H
#define NMSLIB_AMD_HNSW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* distance function codes of the optimized index (include/method/hnsw.h:50-58) */
enum { ORA_L2SQR16 = 1, ORA_L2SQR = 2, ORA_NORMCOS = 3, ORA_NEGDOT = 4, ORA_L1 = 5, ORA_LINF = 6 };
/* algoType (src/method/hnsw.cc:491-501) */
enum { ORA_ALGO_HYBRID = 0, ORA_ALGO_V1MERGE = 1, ORA_ALGO_OLD = 2 };

typedef struct {
  int64_t expansions;     /* level-0 candidates expanded */
  int64_t dist_level0;    /* level-0 distance evaluations */
  int64_t dist_upper;     /* distance evaluations in the greedy descent (incl. entry point) */
  int64_t list_ints_read; /* int32 words of adjacency lists read (count + ids) */
} ora_stats;

const char* ora_last_error(void);

void* ora_load(const char* path);
void ora_free(void* h);
/* header fields: n, dim, maxM, maxM0, maxlevel, enterpoint, dist_func_type, memoryPerObject */
void ora_info(void* h, int64_t* out8);
/* raw views (valid while the handle lives) */
const char* ora_level0(void* h);
int32_t ora_level_of(void* h, int64_t i);
const int32_t* ora_upper_list(void* h, int64_t i, int level);

/* Search nq queries (row-major [nq][dim], copied; cosine normalises the copy).
 * Outputs row-major [nq][k] ascending, counts[q] results; unused slots id -1.
 * stats may be NULL; when given it accumulates over all queries. */
int ora_search(void* h, const float* q, int64_t nq, int k, int ef, int algo, int num_threads,
               int32_t* ids, float* dists, int32_t* counts, ora_stats* stats);

/* Model of the product's beam mode (no reference counterpart): FP32 vectors,
 * `beam` expansions per step, array ordered by (distance, internal id).
 * Same output layout as ora_search. */
int ora_search_beam(void* h, const float* q, int64_t nq, int k, int ef, int beam, int num_threads,
                    int32_t* ids, float* dists, int32_t* counts, ora_stats* stats);

/* The beam model's steps for one query: per step 6 ints -- array size before
 * the step, picks, fresh neighbours, admitted items, and the position the
 * smallest item of each of the (at most two) 256-item merge passes takes
 * (-1: no such pass).  Fills up to cap steps; returns their number (-1 on
 * error). */
int64_t ora_beam_trace(void* h, const float* q, int k, int ef, int beam, int32_t* out,
                       int64_t cap);

float ora_dist(int func, const float* q, const float* x, int64_t dim);
void ora_normalize(float* v, int64_t dim);

/* Exact kNN in FP64 over raw data; space: 1 l2 (squared), 2 cosine (1 - cos on raw
 * vectors), 3 negdot, 4 l1, 5 linf.  Ranked by (distance, index). */
int ora_brute_force(const float* data, int64_t n, int64_t dim, const float* q, int64_t nq, int k,
                    int space, int num_threads, int64_t* ids, double* dists);

#ifdef __cplusplus
}
#endif
#endif

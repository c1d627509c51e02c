/*
 * nmslib_amd -- C ABI of the MI355X HNSW query path.
 *
 * One handle = one index object of python_bindings/nmslib.cc's
 * IndexWrapper<float> (nmslib.cc:68-376) restricted to method "hnsw" on the
 * dense float spaces "l2", "cosinesimil", "negdotprod", "l1" and "linf".  Every entry point
 * returns 0 on success or a negative status; the message of the last failure
 * on the calling thread is nmsamd_last_error().  No C++ exception crosses the
 * ABI and the library keeps no caller pointer past a call.  Results follow the
 * reference exactly: ids are the object ids stored in the index, distances
 * are what the optimized search reports (squared L2 for "l2",
 * 1 - cos for "cosinesimil", -dot for "negdotprod", sum |q - x| for "l1",
 * max |q - x| for "linf"), ascending.
 */
#ifndef NMSLIB_AMD_H
#define NMSLIB_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nmsamd_index nmsamd_index;

/* status codes */
#define NMSAMD_OK 0
#define NMSAMD_EINVAL (-1)   /* invalid argument   -> Python ValueError   */
#define NMSAMD_ERUNTIME (-2) /* runtime failure    -> Python RuntimeError */
#define NMSAMD_EIO (-3)      /* file i/o           -> Python RuntimeError */
#define NMSAMD_EDEVICE (-4)  /* HIP device error   -> Python RuntimeError */

/* Message of the last failure on this thread ("" if none). */
const char* nmsamd_last_error(void);

/* nmslib.init(space, space_params, method, data_type, dtype)
 * (python_bindings/nmslib.cc:470-510).  Only method "hnsw", dense float data;
 * an unknown space raises the reference's spacefactory message. */
int nmsamd_init(const char* space, const char* method, nmsamd_index** out);

/* Releases the handle and its device memory. */
int nmsamd_free(nmsamd_index* h);

/* IndexWrapper::loadIndex(filename, load_data) (nmslib.cc:98-113,
 * Hnsw::LoadIndex hnsw.cc:993-1074): reads the index file, uploads the graph
 * and vectors to HBM (on the first search when no device is visible yet),
 * resets query-time params (ef = 20).  load_data first replaces the data
 * points with filename + ".dat" (space.cc:61-86); without it the points added
 * so far stay.  Both formats load: the optimized one, and the regular one
 * (hnsw.cc:943-991), which holds no vectors and is laid over the data points
 * (their number must match) and searched with the space's own distance. */
int nmsamd_load_index(nmsamd_index* h, const char* filename, int load_data);

/* IndexWrapper::saveIndex(filename, save_data) (nmslib.cc:115-125): writes
 * the index byte-for-byte as Hnsw::SaveOptimizedIndex (hnsw.cc:776-806) or,
 * for a regular index, SaveRegularIndexBin (hnsw.cc:810-841) does; save_data
 * also writes the data points to filename + ".dat". */
int nmsamd_save_index(nmsamd_index* h, const char* filename, int save_data);

/* IndexWrapper::addDataPointBatch(data, ids) (nmslib.cc:342-344,
 * readObjectVector :226-258): appends n dense float points of dimension dim
 * ([n][dim] row-major) to the handle's data, as Objects with id ids[i] (or i,
 * the row number within this batch, when ids is NULL) and label -1.
 * *out_count (optional) receives n, the reference's return value. */
int nmsamd_add_data_point_batch(nmsamd_index* h, const float* data, int64_t n, int64_t dim,
                                const int32_t* ids, int64_t* out_count);

/* IndexWrapper::at(pos) -- __getitem__ (nmslib.cc:348, writeObject :308-320):
 * copies up to cap floats of data point `pos` to out (out may be NULL) and
 * its dimension to *dim.  A position outside the data fails with
 * NMSAMD_EINVAL (the reference's std::out_of_range). */
int nmsamd_get_data_point(nmsamd_index* h, int64_t pos, float* out, int64_t cap, int64_t* dim);

/* IndexWrapper::getDistance(pos1, pos2) (nmslib.cc:350-353): the space's
 * IndexTimeDistance between two data points -- for "l2" the square root of
 * the squared distance, for "cosinesimil" 1 - cos of the raw vectors, in the
 * reference's SSE summation order. */
int nmsamd_get_distance(nmsamd_index* h, int64_t pos1, int64_t pos2, float* out);

/* The space's IndexTimeDistance between two vectors of dimension dim (what
 * the reference's query server's getDistance returns,
 * query_server/cpp_client_server/QueryService_server.cpp:276-312): the same
 * functions as nmsamd_get_distance, over caller vectors. */
int nmsamd_vector_distance(nmsamd_index* h, const float* a, const float* b, int64_t dim,
                           float* out);

/* IndexWrapper::createIndex(index_params, print_progress) (nmslib.cc:89-96 ->
 * Hnsw::CreateIndex, hnsw.cc:183-470): builds the HNSW graph over the data
 * added so far on the host CPU and keeps it as the optimized index, ready for
 * saveIndex and for searching (it reaches the device on the first search).
 * n strings "key=value": M (16), efConstruction (200), maxM (M), maxM0 (2M),
 * mult (1/ln M), indexThreadQty (hardware threads), delaunay_type (2; 0-3,
 * 3 with indexThreadQty=1 only), post (0, 1, 2), skip_optimized_index (1 keeps the regular
 * index, searched with the space's own distance and saved in the regular
 * format), searchMethod (ignored); unknown keys raise "Unknown parameters
 * found!".  maxM0 above 64 is rejected (the device searches lists of at most
 * 64).  With indexThreadQty=1 the graph equals the reference's: the
 * level draws continue the calling thread's generator, seeded with 0 on each
 * thread's first use, as the reference's do.  Resets query-time params
 * (ef = 20).  print_progress reports the points inserted to stderr every 5 %. */
int nmsamd_create_index(nmsamd_index* h, const char* const* kv, int n, int print_progress);

/* setQueryTimeParams (nmslib.cc:599-607 -> Hnsw::SetQueryTimeParams,
 * hnsw.cc:472-507): n strings "key=value"; keys ef | efSearch, algoType
 * (hybrid | old | v1merge), searchMethod (ignored); unknown keys raise
 * "Unknown parameters found!".  NULL/0 resets the defaults.  Not to be called
 * while a search on the same handle is in flight. */
int nmsamd_set_query_time_params(nmsamd_index* h, const char* const* kv, int n);

/* IndexWrapper::knnQueryBatch(queries, k, num_threads) (nmslib.cc:142-184)
 * over host memory: q is [nq][dim] row-major float32; ids / dists are
 * [nq][k] row-major (ascending, -1 / 0 padding past counts[i]); counts[nq].
 * num_threads is accepted for compatibility and ignored (the batch runs on
 * the device).  Blocks until the results are in host memory. */
int nmsamd_knn_query_batch(nmsamd_index* h, const float* q, int64_t nq, int64_t dim, int32_t k,
                           int32_t num_threads, int32_t* ids, float* dists, int32_t* counts);

/* Same over device memory already resident in HBM; launches on `stream`
 * (a hipStream_t, NULL = the library's own stream) and returns without
 * waiting for completion.  The handle's launches share its work counter,
 * visited marks and events, so a later search or brute force of the same
 * handle on another stream waits on the device for this one to finish
 * (hipStreamWaitEvent); launches on one stream are ordered by the stream. */
int nmsamd_knn_query_batch_device(nmsamd_index* h, const float* d_q, int64_t nq, int32_t k,
                                  int32_t* d_ids, float* d_dists, int32_t* d_counts,
                                  void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same search as nmsamd_knn_query_batch_device, same
 * arguments and results, but it calls no hipMalloc / hipFree: its scratch
 * (work counter, error word, epochs, visited marks, SearchOld heaps) is a
 * file-scope device array of 256 MiB per GPU, held until the process ends.
 * As many queries run at once as visited slots of n + 1 bytes (plus the
 * SearchOld heap when that search runs) fit in it; an index too large for
 * one slot is refused.  The index must already be in device memory, and a
 * BF16 / FP8 traversal's converted vectors must exist (a search through
 * another entry point makes both).  No counters are collected.  Every call
 * on one GPU shares its array and clears what it uses first: calls that
 * overlap in time on one GPU race. */
int nmsamd_knn_query_batch_device_noalloc(nmsamd_index* h, const float* d_q, int64_t nq,
                                          int32_t k, int32_t* d_ids, float* d_dists,
                                          int32_t* d_counts, void* stream);

/* Uneven batch: per-query k[i] and ef[i] (host arrays) in one persistent
 * launch (no reference counterpart: the reference runs one ef per index and
 * one k per call).  Output rows have stride max_k = max(k[i]). */
int nmsamd_knn_query_batch_uneven(nmsamd_index* h, const float* q, int64_t nq, int64_t dim,
                                  const int32_t* k, const int32_t* ef, int32_t max_k,
                                  int32_t* ids, float* dists, int32_t* counts);

/* __len__ (nmslib.cc:348): number of data points held with the index
 * (0 when loaded without data, as in the reference). */
int64_t nmsamd_size(const nmsamd_index* h);

/* Index facts: n, dim, maxM, maxM0, maxlevel, enterpoint, dist_func_type,
 * current ef. */
int nmsamd_index_info(const nmsamd_index* h, int64_t* out8);

/* Search precision: 0 = FP32 vectors (reference arithmetic), 1 = BF16,
 * 2 = FP8 e4m3 with a per-vector scale; rerank_mult > 0 reranks the best
 * rerank_mult*k (at least k, at most the traversal capacity) candidates
 * exactly in FP32.  No reference counterpart (extension of the query path). */
int nmsamd_set_precision(nmsamd_index* h, int storage, int rerank_mult);

/* Search mode: 0 = the reference's SearchV1Merge replayed exactly (default);
 * 1..64 = beam mode, that many unexpanded items expanded per step (at most
 * 512 / maxM0, or 2048 / maxM0 on the cooperative kernel, which widths above
 * 16 take) with ties ordered by internal id (no reference counterpart; same
 * array, capacity and stop rule, fewer serial steps).  Beam mode runs every
 * query of a batch, whatever its ef (the hybrid switch to SearchOld at ef >=
 * 1000 does not apply); algoType=old still runs SearchOld. */
int nmsamd_set_beam(nmsamd_index* h, int beam);

/* Waves per query in beam mode: 1, 2 (a block of two waves per query: the
 * second computes each step's distances while the first rebuilds the sorted
 * array from the last merge; at most three queries per CU), 8 (the
 * cooperative kernel: a block of eight waves per query sharing every phase
 * of a step, two queries per CU where LDS holds two, else one; also taken by
 * beam widths above 16), or 0 (default) = two where the LDS footprint limits a
 * CU to three queries.  Results do not depend on it.  No reference
 * counterpart. */
int nmsamd_set_beam_waves(nmsamd_index* h, int waves);

/* Vector loads in flight in the one-wave beam kernels: 1 = the plain kernel
 * (two waves per SIMD; one FP32 / two BF16 / four FP8 vectors per lane pair
 * and round), 2 = the wide one (one wave per SIMD; two FP32 / five BF16 / eight
 * FP8 vectors per lane pair, at most four waves per CU), 0 (default) = wide
 * where the LDS footprint already holds a CU to four waves.  On the
 * cooperative kernel: 2 = one block per CU on the wide path, else two blocks
 * per CU (one FP32 / two BF16 / four FP8 vectors per lane pair) where LDS
 * holds two queries.  Results do not depend on it.  No reference counterpart
 * (an autotuner knob). */
int nmsamd_set_beam_gather(nmsamd_index* h, int gather);

/* BF16 screen of the cooperative FP32 beam search on squared l2 (an optimized
 * index, FP32 storage, setBeamWaves(8) or beam above 16): 1 = with the array
 * full, each candidate's distance is first bounded from below using the BF16
 * copy of its vector (a bound proven to hold for the FP32 distance in the
 * reference's summation order), and only candidates whose bound is below the
 * admission key load their FP32 vector; the admitted items, their FP32
 * distances and every result are those of the search without the screen.
 * 2 = the same, and every candidate's FP32 distance is also computed to check
 * the bound (a failure is reported as an error).  0 (default) = off.  The first
 * such search makes the BF16 copy (2 bytes per coordinate).  Ignored by the
 * other kernels.  No reference counterpart. */
int nmsamd_set_screen(nmsamd_index* h, int mode);

/* The last search's screen counters (with statistics collection on): out2[0] =
 * candidates screened, out2[1] = the ones that passed to their FP32 distance. */
int nmsamd_last_screen_stats(const nmsamd_index* h, uint64_t* out2);

/* Visited marks of the beam-mode searches (the reference keeps a byte per
 * node with a version counter, hnsw.h:568-591): 1 = epoch bytes (n bytes per
 * resident query, cleared when the epoch wraps), 2 = bits (n / 8 bytes,
 * cleared when each query starts), 0 (default) = bits where clearing n / 8
 * bytes is small next to about efSearch * maxM0 mark probes per query.  The
 * exact searches always take epoch bytes.  Results do not depend on it.
 * Servers choose at open by their max_ef.  No reference counterpart (a tuning
 * knob). */
int nmsamd_set_visited_marks(nmsamd_index* h, int marks);

/* Per-call counters of the last search (expansions, level-0 distance
 * evaluations, upper-level distance evaluations, adjacency ints read,
 * expansions that took the tie path) and the kernel time of the last search
 * in milliseconds measured with HIP events on the launch stream. */
int nmsamd_last_stats(const nmsamd_index* h, uint64_t* out5, double* kernel_ms);

/* Turns counter collection on (1) or off (0); off by default.  When on, the
 * device also records each query's search time. */
int nmsamd_collect_stats(nmsamd_index* h, int on);

/* Per-query device search time of the last search with collection on, in
 * microseconds (first nq queries).  No reference counterpart (the
 * reference times queries on the host, experiments.h:186-194). */
int nmsamd_last_query_us(const nmsamd_index* h, float* out_us, int64_t nq);

/* Beam mode, collection on: device time of the last search per phase of a
 * step (pick, adjacency + visited, distances, filter, sort + next-pick
 * prediction, rank search, rebuild of the array, advance), in microseconds
 * summed over queries.  Diagnostic; no reference counterpart. */
int nmsamd_last_phase_us(const nmsamd_index* h, double* out8);

/* Shape of the last search launch: kernel variant (0 = hnsw_search_kernel
 * with V1Merge / SearchOld, 1 = hnsw_search_kernel beam search, 2 =
 * hnsw_search_kernel_wide beam search, 3 = hnsw_search_kernel_pair beam search
 * with two waves per query), blocks, queries per block, LDS bytes per query,
 * neighbour entries per step.  Diagnostic; no reference counterpart. */
int nmsamd_last_launch(const nmsamd_index* h, int64_t* out5);

/* Exact k-NN over the loaded index's FP32 vectors: the device counterpart of
 * the reference's seq_search method (method/seqsearch.cc:143-177) and of its
 * gold standard (gold_standard.h:151-175: ranked by (distance, id)).  A
 * matrix-core screen keeps the best kc = min(256, max(k + 32, 2k)) candidates
 * of each query; their distances are recomputed in FP32 in the reference's
 * summation order and ranked by (distance, internal id).  A query whose
 * screen could have missed an answer (its kc-th screen distance does not
 * clear the k-th exact one by twice the screen's error bound) is searched
 * again by an exact scan of every point, so the result is exact.  Distances
 * are those the HNSW search reports (squared L2, 1 - cos, -dot, AVX order);
 * optimized l2 / cosinesimil / negdotprod indices, dims <= 128,
 * 1 <= k <= 128.  Output [nq][k] rows, counts[nq]. */
int nmsamd_brute_force(nmsamd_index* h, const float* q, int64_t nq, int64_t dim, int32_t k,
                       int32_t* ids, float* dists, int32_t* counts);

/* SeqSearch::Search (method/seqsearch.cc:143-177) over the index's stored
 * vectors: the same exact search ranked and reported in the space's own
 * distance as the reference's seq_search computes it -- l2: L2NormSIMD (the
 * square root of the SSE-order sum, distcomp_lp.cc:304-371); cosinesimil:
 * CosineSimilarity (distcomp_scalar.cc:267-271) of the stored (normalised)
 * vectors; negdotprod: -ScalarProductSIMD -- ties by internal id. */
int nmsamd_seq_search(nmsamd_index* h, const float* q, int64_t nq, int64_t dim, int32_t k,
                      int32_t* ids, float* dists, int32_t* counts);

/* No counterpart in the reference (diagnostic): how many queries of the last
 * brute force / seq_search needed the exact scan of every point. */
int nmsamd_last_bruteforce_rescans(const nmsamd_index* h, int64_t* out);

/* ---- continuous-batching server ------------------------------------------
 * No reference counterpart in the library; it stands where the reference's
 * Thrift query server (query_server/cpp_client_server/QueryService_server.cpp:
 * 411-466) answers knnQueryBatch requests.  A server owns `lanes` independent
 * launch slots over a loaded index (own stream, pinned staging, device
 * buffers and visited-mark slots), so batches submitted to different lanes
 * overlap on the device.  Each batch carries per-query k and ef (1 <= k <=
 * max_k, 0 <= ef <= max_ef) and runs V1Merge (or the index's beam mode /
 * precision settings at open time).  use_graph captures each lane's launch
 * as a hipGraph once and replays it per batch.  The index handle must not be
 * reconfigured while a server is open on it: nmsamd_load_index,
 * nmsamd_create_index and nmsamd_free fail with NMSAMD_EINVAL until its
 * servers are closed. */
typedef struct nmsamd_server nmsamd_server;

int nmsamd_server_open(nmsamd_index* h, int lanes, int max_queries, int max_ef, int max_k,
                       int use_graph, nmsamd_server** out);
/* Copies the batch (q [nq][dim], k[nq], ef[nq]) and launches it on `lane`;
 * returns at once.  The lane must be idle. */
int nmsamd_server_submit(nmsamd_server* s, int lane, const float* q, int64_t nq, const int32_t* k,
                         const int32_t* ef);
/* 1 when the lane's batch is complete (or the lane is idle), 0 while running. */
int nmsamd_server_poll(nmsamd_server* s, int lane);
/* Copies the completed batch's results out ([nq][max_k] rows, counts[nq])
 * and frees the lane. */
int nmsamd_server_fetch(nmsamd_server* s, int lane, int32_t* ids, float* dists, int32_t* counts);
/* lanes, max_queries, max_ef, max_k, slots per lane, waves per block, graph,
 * beam kernel (0 one wave, 1 one wave wide, 2 two waves, 3 cooperative, 4 cooperative at two
 * blocks per CU) */
int nmsamd_server_info(const nmsamd_server* s, int64_t* out8);
int nmsamd_server_close(nmsamd_server* s);

#ifdef __cplusplus
}
#endif
#endif

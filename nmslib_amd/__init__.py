"""nmslib_amd -- the nmslib HNSW query path on MI355X.

A drop-in for the dense-float HNSW subset of the ``nmslib`` Python module
(python_bindings/nmslib.cc): ``init`` returns an index whose ``loadIndex``
reads an optimized index file the reference built and saved, whose
``setQueryTimeParams`` takes ``ef`` / ``efSearch`` / ``algoType`` and whose
``knnQuery`` / ``knnQueryBatch`` answer on the GPU through the C ABI of
``include/nmslib_amd.h`` (library ``nmslib_amd/libnmslib_amd.so``).

    import nmslib_amd as nmslib
    index = nmslib.init(method="hnsw", space="l2")
    index.loadIndex("index.bin")
    index.setQueryTimeParams({"efSearch": 100})
    results = index.knnQueryBatch(queries, k=10)   # [(ids, dists), ...]

Errors map as in the reference binding: invalid arguments raise
``ValueError``, everything else ``RuntimeError``.  The library must be
present: nothing here falls back to a CPU search.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

__all__ = ["init", "DataType", "DistType", "FloatIndex", "library_path", "load_library"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libnmslib_amd.so")

_EINVAL = -1


class DistType(enum.IntEnum):
    FLOAT = 0
    INT = 1


class DataType(enum.IntEnum):
    DENSE_VECTOR = 0
    DENSE_UINT8_VECTOR = 1
    SPARSE_VECTOR = 2
    OBJECT_AS_STRING = 3


_lib = None

# name -> (restype, argtypes) of every entry point of include/nmslib_amd.h
_P, _I64, _I32 = C.c_void_p, C.c_int64, C.c_int32
SIGNATURES = {
    "nmsamd_last_error": (C.c_char_p, []),
    "nmsamd_init": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "nmsamd_free": (C.c_int, [_P]),
    "nmsamd_load_index": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "nmsamd_save_index": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "nmsamd_add_data_point_batch": (C.c_int, [_P, _P, _I64, _I64, _P, C.POINTER(C.c_int64)]),
    "nmsamd_create_index": (C.c_int, [_P, C.POINTER(C.c_char_p), C.c_int, C.c_int]),
    "nmsamd_get_data_point": (C.c_int, [_P, _I64, _P, _I64, C.POINTER(C.c_int64)]),
    "nmsamd_get_distance": (C.c_int, [_P, _I64, _I64, C.POINTER(C.c_float)]),
    "nmsamd_set_query_time_params": (C.c_int, [_P, C.POINTER(C.c_char_p), C.c_int]),
    "nmsamd_knn_query_batch": (C.c_int, [_P, _P, _I64, _I64, _I32, _I32, _P, _P, _P]),
    "nmsamd_knn_query_batch_device": (C.c_int, [_P, _P, _I64, _I32, _P, _P, _P, _P]),
    "nmsamd_knn_query_batch_device_noalloc": (C.c_int, [_P, _P, _I64, _I32, _P, _P, _P, _P]),
    "nmsamd_seq_search": (C.c_int, [_P, _P, _I64, _I64, _I32, _P, _P, _P]),
    "nmsamd_vector_distance": (C.c_int, [_P, _P, _P, _I64, C.POINTER(C.c_float)]),
    "nmsamd_last_bruteforce_rescans": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "nmsamd_knn_query_batch_uneven": (C.c_int, [_P, _P, _I64, _I64, _P, _P, _I32, _P, _P, _P]),
    "nmsamd_size": (_I64, [_P]),
    "nmsamd_index_info": (C.c_int, [_P, _P]),
    "nmsamd_set_precision": (C.c_int, [_P, C.c_int, C.c_int]),
    "nmsamd_last_stats": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "nmsamd_collect_stats": (C.c_int, [_P, C.c_int]),
    "nmsamd_last_query_us": (C.c_int, [_P, _P, _I64]),
    "nmsamd_set_beam": (C.c_int, [_P, C.c_int]),
    "nmsamd_set_beam_gather": (C.c_int, [_P, C.c_int]),
    "nmsamd_set_visited_marks": (C.c_int, [_P, C.c_int]),
    "nmsamd_set_screen": (C.c_int, [_P, C.c_int]),
    "nmsamd_last_screen_stats": (C.c_int, [_P, _P]),
    "nmsamd_set_beam_waves": (C.c_int, [_P, C.c_int]),
    "nmsamd_brute_force": (C.c_int, [_P, _P, _I64, _I64, _I32, _P, _P, _P]),
    "nmsamd_last_phase_us": (C.c_int, [_P, _P]),
    "nmsamd_last_launch": (C.c_int, [_P, _P]),
}


def load_library():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError(f"{library_path} is missing: build it with `python -m nmslib_amd.build`")
        lib = C.CDLL(library_path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        msg = load_library().nmsamd_last_error().decode(errors="replace")
        if rc == _EINVAL:
            raise ValueError(msg)
        raise RuntimeError(msg)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _params_to_kv(params):
    """loadParams (nmslib.cc:658-692): None, a list of 'k=v' strings or a dict of
    int / float / str values."""
    if params is None:
        return []
    if isinstance(params, list):
        return [str(p) for p in params]
    if isinstance(params, dict):
        out = []
        for key, value in params.items():
            if isinstance(value, bool) or not isinstance(value, (int, float, str)):
                raise ValueError(f"Unknown type for parameter '{key}'")
            out.append(f"{key}={value}")
        return out
    raise ValueError("Unknown type for parameters")


def _as_queries(x, dim):
    """readObjectVector semantics: any array-like, force-cast to C-order float32."""
    a = np.asarray(x)
    if a.dtype == object:
        a = np.array([np.asarray(v, dtype=np.float32) for v in x])
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.ndim != 2:
        raise ValueError("queries must be a 2-D array or a list of vectors")
    if dim and a.shape[1] != dim:
        raise ValueError(f"query dimension {a.shape[1]} does not match the index dimension {dim}")
    return a


class FloatIndex:
    """IndexWrapper<float> for method 'hnsw' (nmslib.cc:68-376)."""

    def __init__(self, space: str, method: str, data_type: DataType, dtype: DistType):
        self.space = space
        self.method = method
        self.dataType = data_type
        self.distType = dtype
        h = C.c_void_p()
        _check(load_library().nmsamd_init(space.encode(), method.encode(), C.byref(h)))
        self._h = h
        self._dim = 0
        self._algo = "hybrid"  # algoType as last set (hnsw.cc:491-501)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.nmsamd_free(h)
            self._h = None

    # -- persistence ------------------------------------------------------
    def loadIndex(self, filename: str, load_data: bool = False):
        _check(load_library().nmsamd_load_index(self._h, os.fsencode(filename), int(bool(load_data))))
        self._dim = self.info()["dim"]
        self._algo = "hybrid"

    def saveIndex(self, filename: str, save_data: bool = False):
        _check(load_library().nmsamd_save_index(self._h, os.fsencode(filename), int(bool(save_data))))

    def createIndex(self, index_params=None, print_progress=False):
        """Builds the HNSW graph over the added data (nmslib.cc:89-96); the
        parameters are those of Hnsw::CreateIndex (hnsw.cc:187-205)."""
        kv = _params_to_kv(index_params)
        arr = (C.c_char_p * max(1, len(kv)))(*[s.encode() for s in kv])
        _check(load_library().nmsamd_create_index(self._h, arr, len(kv), int(bool(print_progress))))
        self._dim = self.info()["dim"]
        self._algo = "hybrid"

    def addDataPoint(self, id, data):
        """Appends one point with object id `id`; returns its position
        (nmslib.cc:337-340)."""
        v = np.ascontiguousarray(np.asarray(data, dtype=np.float32).reshape(1, -1))
        ids = np.array([id], np.int32)
        _check(load_library().nmsamd_add_data_point_batch(self._h, _ptr(v), 1, v.shape[1], _ptr(ids),
                                                          None))
        return len(self) - 1

    def addDataPointBatch(self, data, ids=None):
        """Appends the rows of `data` (readObjectVector, nmslib.cc:226-258):
        a 2-D array force-cast to C-order float32, or a list of vectors; point
        i gets id ids[i], or i when ids is None.  Returns the number added."""
        if isinstance(data, list):
            a = np.array([np.asarray(v, dtype=np.float32) for v in data], dtype=np.float32)
        else:
            a = np.asarray(data)
            if a.ndim != 2:
                raise RuntimeError("data must be a 2d array")
            a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim != 2:
            raise RuntimeError("data must be a 2d array")
        idp = None
        if ids is not None:
            idv = np.ascontiguousarray(np.asarray(ids), dtype=np.int32)
            if idv.shape != (a.shape[0],):
                raise ValueError("ids must hold one id per data point")
            idp = _ptr(idv)
        out = C.c_int64(0)
        _check(load_library().nmsamd_add_data_point_batch(self._h, _ptr(a), a.shape[0], a.shape[1],
                                                          idp, C.byref(out)))
        return int(out.value)

    # -- query-time parameters -------------------------------------------
    def setQueryTimeParams(self, params=None):
        kv = _params_to_kv(params)
        arr = (C.c_char_p * max(1, len(kv)))(*[s.encode() for s in kv])
        _check(load_library().nmsamd_set_query_time_params(self._h, arr, len(kv)))
        self._algo = "hybrid"
        for p in kv:
            if p.split("=", 1)[0] == "algoType":
                self._algo = p.split("=", 1)[1].lower()

    def vectorDistance(self, a, b) -> float:
        """The space's own distance between two vectors (IndexTimeDistance,
        as the reference's query server's getDistance computes it)."""
        a = np.ascontiguousarray(a, np.float32).ravel()
        b = np.ascontiguousarray(b, np.float32).ravel()
        if a.shape != b.shape:
            raise ValueError("vectors differ in dimension")
        out = C.c_float()
        _check(load_library().nmsamd_vector_distance(self._h, _ptr(a), _ptr(b), a.shape[0],
                                                     C.byref(out)))
        return float(out.value)

    def setPrecision(self, storage: str = "fp32", rerank_mult: int = 0):
        """Vector storage used while traversing ('fp32', 'bf16', 'fp8') and the
        FP32 rerank width as a multiple of k (0 = no rerank)."""
        code = {"fp32": 0, "bf16": 1, "fp8": 2}[storage]
        _check(load_library().nmsamd_set_precision(self._h, code, int(rerank_mult)))

    def setBeam(self, beam: int = 0):
        """0: exact V1Merge (the reference's results); 1..64: beam mode
        (expansions per step, capped at 512 // maxM0, or 2048 // maxM0 on
        the cooperative kernel that widths above 16 take)."""
        _check(load_library().nmsamd_set_beam(self._h, int(beam)))

    def setBeamWaves(self, waves: int = 0):
        """Waves per beam query: 1, 2 (the second wave computes each step's
        distances while the first rebuilds the sorted array from the last
        merge), 8 (the cooperative kernel: every phase of a step shared by a
        block of eight waves, two queries per CU where LDS holds two; the
        fastest and the low-latency form at large ef) or 0 = chosen from the
        LDS footprint.  Results do not depend on it."""
        _check(load_library().nmsamd_set_beam_waves(self._h, int(waves)))

    def setBeamGather(self, gather: int = 0):
        """Vector loads in flight in the one-wave beam kernels: 1 = plain (two
        waves per SIMD), 2 = wide (one wave per SIMD, every load of a round in
        flight), 0 = chosen from the LDS footprint.  On the cooperative kernel,
        2 keeps one block per CU on the wide path; otherwise two blocks share a
        CU where LDS holds two queries.  Results do not depend on it."""
        _check(load_library().nmsamd_set_beam_gather(self._h, int(gather)))

    def setScreen(self, mode: int = 0):
        """BF16 screen of the cooperative FP32 l2 beam search: 1 = candidates
        whose proven lower bound (from their BF16 copy) is at or above the
        admission key skip their FP32 vector, 2 = the same and every bound is
        checked against the FP32 distance, 0 = off.  Results do not depend on
        it."""
        _check(load_library().nmsamd_set_screen(self._h, int(mode)))

    def setVisitedMarks(self, marks: int = 0):
        """Visited marks of beam mode: 1 = epoch bytes, 2 = bits cleared per
        query, 0 = bits where clearing n / 8 bytes is small next to efSearch *
        maxM0 probes (the exact searches keep epoch bytes).  Results do not
        depend on it."""
        _check(load_library().nmsamd_set_visited_marks(self._h, int(marks)))

    def collectStats(self, on: bool = True):
        _check(load_library().nmsamd_collect_stats(self._h, int(bool(on))))

    def lastStats(self):
        out = np.zeros(5, np.uint64)
        ms = C.c_double()
        _check(load_library().nmsamd_last_stats(self._h, _ptr(out), C.byref(ms)))
        keys = ("expansions", "dist_level0", "dist_upper", "list_ints_read", "tie_paths")
        d = {k: int(v) for k, v in zip(keys, out)}
        d["kernel_ms"] = ms.value
        scr = np.zeros(2, np.uint64)
        _check(load_library().nmsamd_last_screen_stats(self._h, _ptr(scr)))
        d["screened"], d["screen_pass"] = int(scr[0]), int(scr[1])
        return d

    def lastPhaseUs(self):
        """With collectStats(True): device time per phase of a step (beam
        mode) or of an expansion (exact V1Merge), us summed over the queries
        of the last search."""
        out = np.zeros(8, np.float64)
        _check(load_library().nmsamd_last_phase_us(self._h, _ptr(out)))
        mode = self.lastLaunch()["mode"]
        if mode == "v1merge/old":
            keys = ("pick_adjacency", "visited", "distances", "admit_predict", "sort",
                    "rank_search", "merge", "advance")
        elif mode.startswith("beam (cooperative"):
            keys = ("pick", "rows_marks", "distances_admit", "rank_new", "rank_array",
                    "records", "rebuild", "advance")
        else:
            keys = ("pick", "adjacency_visited", "distances", "filter", "sort_predict",
                    "rank_search", "rebuild", "advance")
        return {k: float(v) for k, v in zip(keys, out)}

    def lastLaunch(self):
        """Shape of the last search launch (kernel variant, grid, LDS)."""
        out = np.zeros(5, np.int64)
        _check(load_library().nmsamd_last_launch(self._h, _ptr(out)))
        mode = int(out[0])
        kernel = {2: "hnsw_search_kernel_wide", 3: "hnsw_search_kernel_pair",
                  4: "hnsw_search_kernel_coop", 5: "hnsw_search_kernel_coop2"}.get(
                      mode, "hnsw_search_kernel")
        return {"kernel": kernel,
                "mode": ("v1merge/old", "beam", "beam (wide)", "beam (two waves)",
                         "beam (cooperative)", "beam (cooperative, two per CU)")[mode],
                "blocks": int(out[1]),
                "queries_per_block": int(out[2]), "lds_bytes_per_query": int(out[3]),
                "entries_per_step": int(out[4])}

    def lastQueryTimesUs(self, nq: int):
        """Per-query device search time (us) of the last search made with
        collectStats(True)."""
        out = np.empty(nq, np.float32)
        _check(load_library().nmsamd_last_query_us(self._h, _ptr(out), int(nq)))
        return out

    def info(self):
        out = np.zeros(8, np.int64)
        _check(load_library().nmsamd_index_info(self._h, _ptr(out)))
        keys = ("n", "dim", "maxM", "maxM0", "maxlevel", "enterpoint", "dist_func_type", "ef")
        return {k: int(v) for k, v in zip(keys, out)}

    # -- search -------------------------------------------------------------
    def knnQueryBatchArrays(self, queries, k: int = 10):
        """Batch search returning (ids[nq, k], dists[nq, k], counts[nq])."""
        q = _as_queries(queries, self._dim)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        _check(load_library().nmsamd_knn_query_batch(self._h, _ptr(q), nq, q.shape[1], int(k), 0,
                                                     _ptr(ids), _ptr(dists), _ptr(counts)))
        return ids, dists, counts

    def knnQueryBatch(self, queries, k: int = 10, num_threads: int = 0):
        """knnQueryBatch (nmslib.cc:142-184): a list of (ids int32, dists float32)
        per query, ascending.  num_threads is accepted and ignored."""
        ids, dists, counts = self.knnQueryBatchArrays(queries, k)
        # one row view per query of the result arrays (a per-row copy costs
        # about four times as much); rows with fewer than k results are cut
        rows_i, rows_d = list(ids), list(dists)
        for i in np.nonzero(counts < ids.shape[1])[0]:
            rows_i[i] = rows_i[i][:counts[i]]
            rows_d[i] = rows_d[i][:counts[i]]
        return list(zip(rows_i, rows_d))

    def knnQuery(self, vector, k: int = 10):
        ids, dists = self.knnQueryBatch(np.asarray(vector, np.float32).reshape(1, -1), k)[0]
        return ids, dists

    def bruteForce(self, queries, k: int = 10):
        """Exact k-NN on the device (matrix-core screen, exact FP32 re-scoring):
        (ids[nq, k], dists[nq, k], counts[nq]) by (distance, id) in the
        distance the HNSW search reports."""
        q = _as_queries(queries, self._dim)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        _check(load_library().nmsamd_brute_force(self._h, _ptr(q), nq, q.shape[1], int(k),
                                                 _ptr(ids), _ptr(dists), _ptr(counts)))
        return ids, dists, counts

    def seqSearch(self, queries, k: int = 10):
        """The reference's seq_search on the device: exact k-NN in the space's
        own distance (l2 with the square root, SSE summation order), ranked by
        (distance, id): (ids[nq, k], dists[nq, k], counts[nq])."""
        q = _as_queries(queries, self._dim)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        _check(load_library().nmsamd_seq_search(self._h, _ptr(q), nq, q.shape[1], int(k),
                                                _ptr(ids), _ptr(dists), _ptr(counts)))
        return ids, dists, counts

    def lastBruteForceRescans(self) -> int:
        """Queries of the last bruteForce / seqSearch that took the exact scan."""
        out = C.c_int64()
        _check(load_library().nmsamd_last_bruteforce_rescans(self._h, C.byref(out)))
        return int(out.value)

    def knnQueryBatchUneven(self, queries, ks, efs):
        """One persistent launch with per-query k and ef."""
        q = _as_queries(queries, self._dim)
        nq = q.shape[0]
        ks = np.ascontiguousarray(ks, np.int32)
        efs = np.ascontiguousarray(efs, np.int32)
        if ks.shape != (nq,) or efs.shape != (nq,):
            raise ValueError("ks and efs need one entry per query")
        mk = int(ks.max()) if nq else 0
        ids = np.empty((nq, max(mk, 1)), np.int32)
        dists = np.empty((nq, max(mk, 1)), np.float32)
        counts = np.empty(nq, np.int32)
        _check(load_library().nmsamd_knn_query_batch_uneven(
            self._h, _ptr(q), nq, q.shape[1], _ptr(ks), _ptr(efs), max(mk, 1),
            _ptr(ids), _ptr(dists), _ptr(counts)))
        return ids, dists, counts

    def knnQueryBatchDevice(self, q_ptr: int, nq: int, k: int, ids_ptr: int, dists_ptr: int,
                            counts_ptr: int, stream: int = 0):
        """Queries and outputs already in device memory (raw pointers)."""
        _check(load_library().nmsamd_knn_query_batch_device(
            self._h, C.c_void_p(q_ptr), int(nq), int(k), C.c_void_p(ids_ptr),
            C.c_void_p(dists_ptr), C.c_void_p(counts_ptr), C.c_void_p(stream or None)))

    def knnQueryBatchDeviceNoAlloc(self, q_ptr: int, nq: int, k: int, ids_ptr: int,
                                   dists_ptr: int, counts_ptr: int, stream: int = 0):
        """knnQueryBatchDevice without device allocations: the library's
        per-GPU scratch array holds the search state (see the C header)."""
        _check(load_library().nmsamd_knn_query_batch_device_noalloc(
            self._h, C.c_void_p(q_ptr), int(nq), int(k), C.c_void_p(ids_ptr),
            C.c_void_p(dists_ptr), C.c_void_p(counts_ptr), C.c_void_p(stream or None)))

    def __len__(self):
        return int(load_library().nmsamd_size(self._h))

    def __getitem__(self, pos):
        """Data point `pos` as a list of floats (nmslib.cc:348, writeObject
        :308-320); IndexError outside the data, like the reference's .at()."""
        pos = int(pos)
        if pos < 0 or pos >= len(self):
            raise IndexError(f"data point position {pos} out of range")
        dim = C.c_int64(0)
        _check(load_library().nmsamd_get_data_point(self._h, pos, None, 0, C.byref(dim)))
        out = np.empty(dim.value, np.float32)
        _check(load_library().nmsamd_get_data_point(self._h, pos, _ptr(out), dim.value, None))
        return [float(v) for v in out]

    def getDistance(self, pos1, pos2):
        """The space's distance between data points pos1 and pos2
        (nmslib.cc:350-353)."""
        if not (0 <= int(pos1) < len(self) and 0 <= int(pos2) < len(self)):
            raise IndexError("data point position out of range")
        out = C.c_float(0)
        _check(load_library().nmsamd_get_distance(self._h, int(pos1), int(pos2), C.byref(out)))
        return float(out.value)

    def __repr__(self):
        return f"<nmslib_amd.FloatIndex method='{self.method}' space='{self.space}'>"


def init(space: str = "cosinesimil", space_params=None, method: str = "hnsw",
         data_type: DataType = DataType.DENSE_VECTOR, dtype: DistType = DistType.FLOAT):
    """nmslib.init (nmslib.cc:470-510) for dense float HNSW indices."""
    if space_params:
        raise ValueError("space parameters are not used by the l2 / cosinesimil / negdotprod spaces")
    if DataType(data_type) != DataType.DENSE_VECTOR:
        raise ValueError("only DataType.DENSE_VECTOR is supported by the device query path")
    if DistType(dtype) != DistType.FLOAT:
        raise ValueError("only DistType.FLOAT is supported by the device query path")
    return FloatIndex(space, method, DataType(data_type), DistType(dtype))


# -- the legacy module-level API (nmslib.cc:694-789) --------------------------
def addDataPoint(index, id, datum):
    return index.addDataPoint(id, datum)


def addDataPointBatch(index, ids, data):
    """Legacy form: numpy arrays only (ids int32, data float32); returns the
    positions of the new points."""
    if not (isinstance(ids, np.ndarray) and ids.dtype == np.int32):
        raise ValueError("Invalid datatype for ids in addDataPointBatch")
    if not (isinstance(data, np.ndarray) and data.dtype == np.float32):
        raise ValueError("Invalid datatype for data in addDataPointBatch")
    offset = len(index)
    n = index.addDataPointBatch(data, ids)
    return np.arange(offset, offset + n, dtype=np.int32)


def setQueryTimeParams(index, params):
    return index.setQueryTimeParams(params)


def createIndex(index, index_params):
    return index.createIndex(index_params)


def saveIndex(index, filename):
    return index.saveIndex(filename)


def loadIndex(index, filename):
    return index.loadIndex(filename)


def knnQuery(index, k, query):
    """Legacy form: the list of ids only."""
    return [int(i) for i in index.knnQuery(query, k)[0]]


def getDataPoint(index, pos):
    return index[pos]


def getDataPointQty(index):
    return len(index)


def getDistance(index, pos1, pos2):
    return index.getDistance(pos1, pos2)


def knnQueryBatch(index, num_threads, k, queries):
    """Legacy form: a list of id lists."""
    return [[int(i) for i in ids] for ids, _ in index.knnQueryBatch(queries, k, num_threads)]


def freeIndex(index):
    pass

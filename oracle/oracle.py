"""ORACLE -- test infrastructure only.

ctypes front end of ``oracle/liboracle.so``, the CPU restatement of nmslib's
optimized-HNSW query path (reference map in ``oracle/hnsw_oracle.h``).  Only
tests, the smoke check and the CPU-baseline leg of the benchmark import this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

L2SQR16, L2SQR, NORMCOS, NEGDOT, L1, LINF = 1, 2, 3, 4, 5, 6
ALGO = {"hybrid": 0, "v1merge": 1, "old": 2}
SPACE_BF = {"l2": 1, "cosinesimil": 2, "negdotprod": 3, "l1": 4, "linf": 5}


class Stats(C.Structure):
    _fields_ = [("expansions", C.c_int64), ("dist_level0", C.c_int64),
                ("dist_upper", C.c_int64), ("list_ints_read", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        l = C.CDLL(_LIB)
        P, I64, I32 = C.c_void_p, C.c_int64, C.c_int
        l.ora_last_error.restype = C.c_char_p
        l.ora_load.restype = P
        l.ora_load.argtypes = [C.c_char_p]
        l.ora_free.argtypes = [P]
        l.ora_info.argtypes = [P, P]
        l.ora_level0.restype = P
        l.ora_level0.argtypes = [P]
        l.ora_level_of.restype = C.c_int32
        l.ora_level_of.argtypes = [P, I64]
        l.ora_upper_list.restype = P
        l.ora_upper_list.argtypes = [P, I64, I32]
        l.ora_search.restype = I32
        l.ora_search.argtypes = [P, P, I64, I32, I32, I32, I32, P, P, P, P]
        l.ora_search_beam.restype = I32
        l.ora_search_beam.argtypes = [P, P, I64, I32, I32, I32, I32, P, P, P, P]
        l.ora_beam_trace.restype = C.c_int64
        l.ora_beam_trace.argtypes = [P, P, I32, I32, I32, P, I64]
        l.ora_dist.restype = C.c_float
        l.ora_dist.argtypes = [I32, P, P, I64]
        l.ora_normalize.argtypes = [P, I64]
        l.ora_brute_force.restype = I32
        l.ora_brute_force.argtypes = [P, I64, I64, P, I64, I32, I32, I32, P, P]
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleIndex:
    """A loaded reference-format optimized HNSW index, searched on the CPU."""

    def __init__(self, path: str):
        self._h = lib().ora_load(path.encode())
        if not self._h:
            raise RuntimeError(lib().ora_last_error().decode())
        info = np.zeros(8, np.int64)
        lib().ora_info(self._h, _p(info))
        (self.n, self.dim, self.maxM, self.maxM0, self.maxlevel, self.enterpoint,
         self.func, self.elem_bytes) = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_free(self._h)
            self._h = None

    def search(self, q, k: int, ef: int, algo: str = "hybrid", num_threads: int = 1,
               with_stats: bool = False):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1, self.dim)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        st = Stats()
        rc = lib().ora_search(self._h, _p(q), nq, k, ef, ALGO[algo], num_threads,
                              _p(ids), _p(dists), _p(counts), C.byref(st))
        if rc:
            raise RuntimeError(lib().ora_last_error().decode())
        if with_stats:
            return ids, dists, counts, {f: getattr(st, f) for f, _ in Stats._fields_}
        return ids, dists, counts

    def search_beam(self, q, k: int, ef: int, beam: int, num_threads: int = 1,
                    with_stats: bool = False):
        """Model of the product's beam mode (FP32 vectors, `beam` expansions
        per step, ties by internal id); no reference counterpart."""
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1, self.dim)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        st = Stats()
        rc = lib().ora_search_beam(self._h, _p(q), nq, k, ef, beam, num_threads,
                                   _p(ids), _p(dists), _p(counts), C.byref(st))
        if rc:
            raise RuntimeError(lib().ora_last_error().decode())
        if with_stats:
            return ids, dists, counts, {f: getattr(st, f) for f, _ in Stats._fields_}
        return ids, dists, counts

    def beam_trace(self, q, k: int, ef: int, beam: int) -> np.ndarray:
        """Per-step records of the beam model for one query: columns n,
        picks, fresh, admitted, first-item position of merge pass 0 and 1."""
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(self.dim)
        cap = 1 << 16
        out = np.empty((cap, 6), np.int32)
        n = lib().ora_beam_trace(self._h, _p(q), k, ef, beam, _p(out), cap)
        if n < 0:
            raise RuntimeError(lib().ora_last_error().decode())
        return out[:min(n, cap)].copy()

    def level0_block(self) -> np.ndarray:
        ptr = lib().ora_level0(self._h)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)),
                                     shape=(self.n * self.elem_bytes,))


def dist(func: int, q, x) -> float:
    q = np.ascontiguousarray(q, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().ora_dist(func, _p(q), _p(x), q.shape[0]))


def normalize(v) -> np.ndarray:
    v = np.array(v, dtype=np.float32, copy=True)
    lib().ora_normalize(_p(v), v.shape[0])
    return v


def brute_force(data, q, k: int, space: str, num_threads: int = 0):
    data = np.ascontiguousarray(data, np.float32)
    q = np.ascontiguousarray(q, np.float32).reshape(-1, data.shape[1])
    ids = np.empty((q.shape[0], k), np.int64)
    dists = np.empty((q.shape[0], k), np.float64)
    rc = lib().ora_brute_force(_p(data), data.shape[0], data.shape[1], _p(q), q.shape[0], k,
                               SPACE_BF[space], num_threads, _p(ids), _p(dists))
    if rc:
        raise RuntimeError(lib().ora_last_error().decode())
    return ids, dists

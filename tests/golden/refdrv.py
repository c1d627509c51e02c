"""ctypes front end of oracle/_ref/libnmsref.so: the reference nmslib HNSW
sources compiled here by ``make -C oracle ref`` with our own driver
(oracle/ref_driver.cc).  Used only by the golden-fixture generator in this
folder, which runs where the reference sources exist."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
_LIB = os.path.join(_ROOT, "oracle", "_ref", "libnmsref.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        l = C.CDLL(_LIB)
        P, I64, I32 = C.c_void_p, C.c_int64, C.c_int
        l.ref_last_error.restype = C.c_char_p
        l.ref_create.restype = P
        l.ref_create.argtypes = [C.c_char_p, P, I64, I32]
        l.ref_create_index.argtypes = [P, C.c_char_p]
        l.ref_save_index.argtypes = [P, C.c_char_p]
        l.ref_load.restype = P
        l.ref_load.argtypes = [C.c_char_p, C.c_char_p, I32]
        l.ref_set_query_params.argtypes = [P, C.c_char_p]
        l.ref_knn_query_batch.argtypes = [P, P, I64, I32, I32, I32, P, P, P]
        l.ref_seq_search.argtypes = [P, P, I64, I32, I32, P, P, P]
        l.ref_free.argtypes = [P]
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _chk(rc):
    if rc:
        raise RuntimeError(lib().ref_last_error().decode())


class RefIndex:
    def __init__(self, handle, dim):
        if not handle:
            raise RuntimeError(lib().ref_last_error().decode())
        self.h, self.dim = handle, dim

    @classmethod
    def build(cls, space, data, params):
        data = np.ascontiguousarray(data, np.float32)
        r = cls(lib().ref_create(space.encode(), _p(data), data.shape[0], data.shape[1]), data.shape[1])
        _chk(lib().ref_create_index(r.h, params.encode()))
        return r

    @classmethod
    def load(cls, space, path, dim):
        return cls(lib().ref_load(space.encode(), path.encode(), dim), dim)

    def save(self, path):
        _chk(lib().ref_save_index(self.h, path.encode()))

    def set_query_params(self, params):
        _chk(lib().ref_set_query_params(self.h, params.encode()))

    def knn_query_batch(self, q, k, num_threads=1):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        _chk(lib().ref_knn_query_batch(self.h, _p(q), nq, self.dim, k, num_threads,
                                       _p(ids), _p(dists), _p(counts)))
        return ids, dists, counts

    def seq_search(self, q, k):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        dists = np.empty((nq, k), np.float32)
        counts = np.empty(nq, np.int32)
        _chk(lib().ref_seq_search(self.h, _p(q), nq, self.dim, k, _p(ids), _p(dists), _p(counts)))
        return ids, dists, counts

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_free(self.h)
            self.h = None

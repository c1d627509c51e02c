"""Continuous-batching request loop over the device HNSW index.

The reference answers served queries through its Thrift server
(query_server/cpp_client_server/QueryService_server.cpp:411-466: a
knnQueryBatch request runs ParallelFor over Index::Search).  Here a request
(1..256 queries with one k and one ef) is admitted into the next batch of a
free server lane as soon as it arrives; lanes run concurrently on the device
(their own streams and visited slots), so requests that arrive while others
are on the GPU start as soon as a lane frees up.  Admission is bounded by the
lane budget: a batch holds at most ``max_queries`` queries and only requests
whose max(ef, k) fits the LDS footprint the server was opened for.

``Server`` wraps the C ABI (nmsamd_server_*); ``run_poisson`` drives it with a
seeded Poisson arrival process and reports per-request latency (completion
seen by the loop minus arrival) and throughput.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _check, _ptr, load_library

_P, _I64, _I32 = C.c_void_p, C.c_int64, C.c_int32
SERVER_SIGNATURES = {
    "nmsamd_server_open": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_void_p)]),
    "nmsamd_server_submit": (C.c_int, [_P, C.c_int, _P, _I64, _P, _P]),
    "nmsamd_server_poll": (C.c_int, [_P, C.c_int]),
    "nmsamd_server_fetch": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "nmsamd_server_info": (C.c_int, [_P, _P]),
    "nmsamd_server_close": (C.c_int, [_P]),
}


def _lib():
    lib = load_library()
    if not getattr(lib, "_server_sigs", False):
        for name, (res, args) in SERVER_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._server_sigs = True
    return lib


class Server:
    """`lanes` concurrent batch slots over a loaded FloatIndex."""

    def __init__(self, index, lanes: int = 4, max_queries: int = 2048, max_ef: int = 800,
                 max_k: int = 100, use_graph: bool = True):
        self.index = index
        self.dim = index.info()["dim"]
        self.max_queries, self.max_ef, self.max_k = max_queries, max_ef, max_k
        h = C.c_void_p()
        _check(_lib().nmsamd_server_open(index._h, lanes, max_queries, max_ef, max_k,
                                         int(use_graph), C.byref(h)))
        self._h = h
        self.lanes = lanes
        self.busy = [None] * lanes

    def info(self):
        out = np.zeros(8, np.int64)
        _check(_lib().nmsamd_server_info(self._h, _ptr(out)))
        # beam_kernel: 0 = one wave per query (plain), 1 = one wave (wide),
        # 2 = two waves per query, 3 = cooperative (a block of eight waves),
        # 4 = cooperative at two blocks per CU
        keys = ("lanes", "max_queries", "max_ef", "max_k", "slots_per_lane", "waves_per_block",
                "graph", "beam_kernel")
        return {k: int(v) for k, v in zip(keys, out)}

    def submit(self, lane: int, q: np.ndarray, ks: np.ndarray, efs: np.ndarray):
        q = np.ascontiguousarray(q, np.float32)
        ks = np.ascontiguousarray(ks, np.int32)
        efs = np.ascontiguousarray(efs, np.int32)
        _check(_lib().nmsamd_server_submit(self._h, lane, _ptr(q), q.shape[0], _ptr(ks), _ptr(efs)))
        self.busy[lane] = q.shape[0]

    def poll(self, lane: int) -> bool:
        rc = _lib().nmsamd_server_poll(self._h, lane)
        if rc < 0:
            _check(rc)
        return rc == 1

    def fetch(self, lane: int):
        nq = self.busy[lane]
        ids = np.empty((nq, self.max_k), np.int32)
        dists = np.empty((nq, self.max_k), np.float32)
        counts = np.empty(nq, np.int32)
        _check(_lib().nmsamd_server_fetch(self._h, lane, _ptr(ids), _ptr(dists), _ptr(counts)))
        self.busy[lane] = None
        return ids, dists, counts

    def close(self):
        if getattr(self, "_h", None):
            _lib().nmsamd_server_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


@dataclass
class Request:
    rid: int
    arrival: float          # seconds since the start
    rows: np.ndarray        # indices into the query pool
    k: int
    ef: int
    done: float = -1.0
    result: tuple = field(default=None, repr=False)


def make_requests(n_pool: int, rate_qps: float, duration_s: float, seed: int,
                  size_range=(1, 256), ks=(1, 10, 100), ef_range=(50, 800)):
    """Poisson arrivals of requests whose sizes are uniform in size_range, so
    the query arrival rate is rate_qps on average."""
    rng = np.random.default_rng(seed)
    mean_size = (size_range[0] + size_range[1]) / 2
    req_rate = rate_qps / mean_size
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / req_rate)
        if t > duration_s:
            break
        size = int(rng.integers(size_range[0], size_range[1] + 1))
        rows = rng.integers(0, n_pool, size=size)
        out.append(Request(rid, t, rows, int(rng.choice(ks)),
                           int(rng.integers(ef_range[0], ef_range[1] + 1))))
        rid += 1
    return out


def run_poisson(server: Server, pool: np.ndarray, requests, keep_results: bool = False):
    """Serves `requests` in real time; returns latency / throughput stats."""
    pending = []            # arrived, not yet launched
    inflight = [None] * server.lanes   # lane -> list of requests
    nxt = 0
    t0 = time.perf_counter()
    total = len(requests)
    finished = 0
    while finished < total:
        now = time.perf_counter() - t0
        while nxt < total and requests[nxt].arrival <= now:
            pending.append(requests[nxt])
            nxt += 1
        for lane in range(server.lanes):
            batch = inflight[lane]
            if batch is not None and server.poll(lane):
                ids, dists, counts = server.fetch(lane)
                tdone = time.perf_counter() - t0
                off = 0
                for r in batch:
                    n = len(r.rows)
                    r.done = tdone
                    if keep_results:
                        r.result = (ids[off:off + n, :r.k].copy(), dists[off:off + n, :r.k].copy(),
                                    counts[off:off + n].copy())
                    off += n
                finished += len(batch)
                inflight[lane] = None
            if inflight[lane] is None and pending:
                # admit arrived requests while the lane's budget allows
                batch, nq = [], 0
                while pending and nq + len(pending[0].rows) <= server.max_queries:
                    r = pending.pop(0)
                    batch.append(r)
                    nq += len(r.rows)
                q = np.concatenate([pool[r.rows] for r in batch])
                ks = np.concatenate([np.full(len(r.rows), r.k, np.int32) for r in batch])
                efs = np.concatenate([np.full(len(r.rows), r.ef, np.int32) for r in batch])
                server.submit(lane, q, ks, efs)
                inflight[lane] = batch
        if nxt < total and not pending and all(b is None for b in inflight):
            # idle until the next arrival
            delay = requests[nxt].arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(min(delay, 0.001))
    span = max(r.done for r in requests) - requests[0].arrival if requests else 0.0
    # the arrival process's own rate (a Poisson sample of the offered rate)
    arrive = requests[-1].arrival - requests[0].arrival if total > 1 else 0.0
    lat = np.array([r.done - r.arrival for r in requests]) * 1000.0
    nqueries = sum(len(r.rows) for r in requests)
    return {"requests": total, "queries": int(nqueries),
            "qps": nqueries / span if span > 0 else 0.0,
            "arrival_qps": nqueries / arrive if arrive > 0 else 0.0,
            "latency_ms_p50": float(np.percentile(lat, 50)) if total else 0.0,
            "latency_ms_p99": float(np.percentile(lat, 99)) if total else 0.0,
            "latency_ms_max": float(lat.max()) if total else 0.0}

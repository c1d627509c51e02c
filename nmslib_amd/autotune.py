"""Picks the search form for a workload by measurement.

The knobs of the device search are the beam width (expansions per step:
0 = the reference's exact V1Merge, 1..4 = beam mode; with the gathers of a
step in flight together when the LDS footprint already holds residency to
about a wave per SIMD) and the vector storage used while traversing (FP32,
BF16, FP8-e4m3 with a per-vector scale, each followed by an exact FP32 rerank
of the best 2k candidates).  ``autotune`` runs a calibration batch -- with its
own per-query k and ef -- in every candidate form, computes each form's
recall@k against the device brute force, and keeps the fastest form whose
recall is within ``recall_slack`` of the exact form's.  Choices are cached
per (dimension, ef bucket) and applied to the index.
"""
from __future__ import annotations

import time

import numpy as np

DEFAULT_FORMS = [
    # (name, storage, rerank multiple of k, beam width)
    ("fp32-exact", "fp32", 0, 0),
    ("fp32-beam2", "fp32", 0, 2),
    ("fp32-beam4", "fp32", 0, 4),
    ("bf16-beam4-rerank", "bf16", 2, 4),
    ("fp8-beam4-rerank", "fp8", 2, 4),
]

_cache: dict = {}


def _recall(ids, gt_ids, gt_d, ks):
    tot = 0.0
    for a, g, gd, k in zip(ids, gt_ids, gt_d, ks):
        kth = gd[k - 1]
        good = set(int(x) for x, d in zip(g, gd) if d <= kth * (1 + 1e-6) + 1e-12)
        tot += sum(1 for x in a[:k] if int(x) in good) / k
    return tot / len(ids)


def _ef_bucket(efs):
    return int(2 ** np.ceil(np.log2(max(2, int(np.max(efs))))))


def autotune(index, queries, ks, efs, forms=None, recall_slack: float = 0.003, reps: int = 2,
             apply: bool = True):
    queries = np.ascontiguousarray(queries, np.float32)
    ks = np.ascontiguousarray(ks, np.int32)
    efs = np.ascontiguousarray(efs, np.int32)
    key = (queries.shape[1], _ef_bucket(efs), int(ks.max()))
    if key in _cache and apply:
        best = _cache[key]
        index.setPrecision(best["storage"], best["rerank"])
        index.setBeam(best["beam"])
        return best
    forms = forms or DEFAULT_FORMS
    kmax = int(ks.max())
    gt_ids, gt_d, _ = index.bruteForce(queries, min(128, max(kmax, 10)))
    index.setQueryTimeParams({"algoType": "v1merge"})
    report = []
    for name, storage, rerank, beam in forms:
        index.setPrecision(storage, rerank)
        index.setBeam(beam)
        ids, _, _ = index.knnQueryBatchUneven(queries, ks, efs)  # warm: conversions, code objects
        t = time.perf_counter()
        for _ in range(reps):
            ids, _, _ = index.knnQueryBatchUneven(queries, ks, efs)
        el = (time.perf_counter() - t) / reps
        report.append({"form": name, "storage": storage, "rerank": rerank, "beam": beam,
                       "seconds": el, "qps": len(queries) / el,
                       "recall": _recall(ids, gt_ids, gt_d, ks)})
    exact = next((r for r in report if r["beam"] == 0 and r["storage"] == "fp32"), report[0])
    ok = [r for r in report if r["recall"] >= exact["recall"] - recall_slack]
    best = dict(max(ok, key=lambda r: r["qps"]))
    best["candidates"] = report
    best["key"] = key
    if apply:
        index.setPrecision(best["storage"], best["rerank"])
        index.setBeam(best["beam"])
        _cache[key] = best
    else:
        index.setPrecision("fp32", 0)
        index.setBeam(0)
    return best

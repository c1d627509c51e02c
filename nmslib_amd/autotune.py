"""Picks the search form for a workload by measurement.

The knobs of the device search (none changes which results a form returns,
except storage and rerank, whose recall is measured):

* beam width W: expansions per step, 0 = the reference's exact V1Merge,
  1..64 = beam mode;
* waves per query in beam mode: 1, 2 (the second wave computes each step's
  distances while the first rebuilds the sorted array) or 8 (the
  cooperative kernel: a block of eight waves per query, every phase of a step
  shared, two blocks per CU where LDS holds two queries; the only kernel for
  widths above 16);
* gather width of the one-wave kernels: plain (two waves per SIMD, one FP32 /
  two BF16 / four FP8 vectors per lane pair in flight) or wide (one wave per
  SIMD, two / five / eight);
* vector storage while traversing: FP32, BF16 or FP8-e4m3 with a per-vector
  scale, the reduced ones followed by an exact FP32 rerank of the best
  rerank * k candidates (the rerank width);
* the BF16 screen of the cooperative FP32 l2 search (results unchanged).

``autotune`` runs a calibration batch -- with its own per-query k and ef --
in a staged search: every storage x beam width with the kernel chosen from
the LDS footprint, then, for the fastest forms within the recall slack, the
waves per query, the gather width and (reduced precision) the rerank width.
Recall@k is computed against the device brute force (exact), and the fastest
form whose recall is within ``recall_slack`` of the exact FP32 form's is kept.
Choices are cached per (dimension, ef bucket, max k) and applied to the index.
"""
from __future__ import annotations

import time

import numpy as np

# (name, storage, rerank multiple of k, beam width, waves per query, gather,
# BF16 screen)
Form = tuple

STAGE1 = [
    ("fp32-exact", "fp32", 0, 0, 0, 0),
    ("fp32-beam2", "fp32", 0, 2, 0, 0),
    ("fp32-beam4", "fp32", 0, 4, 0, 0),
    ("fp32-beam8", "fp32", 0, 8, 0, 0),
    ("fp32-beam16", "fp32", 0, 16, 0, 0),
    ("bf16-beam4-rerank2", "bf16", 2, 4, 0, 0),
    ("bf16-beam8-rerank2", "bf16", 2, 8, 0, 0),
    ("bf16-beam16-rerank2", "bf16", 2, 16, 0, 0),
    ("fp8-beam8-rerank2", "fp8", 2, 8, 0, 0),
    ("fp32-coop32", "fp32", 0, 32, 8, 0),
    ("fp32-coop48", "fp32", 0, 48, 8, 0),
    ("fp32-coop64-screen", "fp32", 0, 64, 8, 0, 1),
    ("bf16-coop32-rerank2", "bf16", 2, 32, 8, 0),
    ("bf16-coop48-rerank2", "bf16", 2, 48, 8, 0),
    ("bf16-coop64-rerank2", "bf16", 2, 64, 8, 0),
]
# kept for callers that pass the old five forms explicitly
DEFAULT_FORMS = [f[:4] for f in STAGE1]

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


def _apply(index, form):
    _, storage, rerank, beam, waves, gather, screen = (tuple(form) + (0, 0, 0))[:7]
    index.setPrecision(storage, rerank)
    index.setBeam(beam)
    index.setBeamWaves(waves)
    index.setBeamGather(gather)
    index.setScreen(screen)


def _variants(form):
    """Stage 2: the same form with explicit waves / gather / rerank width."""
    name, storage, rerank, beam, waves = form[:5]
    out = []
    if 0 < beam <= 16 and waves != 8:
        out.append((f"{name}-1wave-plain", storage, rerank, beam, 1, 1))
        out.append((f"{name}-1wave-wide", storage, rerank, beam, 1, 2))
        out.append((f"{name}-2waves", storage, rerank, beam, 2, 0))
        out.append((f"{name}-coop", storage, rerank, beam, 8, 0))
    if waves == 8 or beam > 16:  # one cooperative block per CU, wide distance path
        out.append((f"{name}-coop-wide", storage, rerank, beam, 8, 2,
                    form[6] if len(form) > 6 else 0))
    if storage != "fp32":
        for r in (1, 4):
            out.append((f"{name.rsplit('-rerank', 1)[0]}-rerank{r}", storage, r, beam, waves, 0))
    return out


def autotune(index, queries, ks, efs, forms=None, recall_slack: float = 0.003, reps: int = 2,
             apply: bool = True, refine: int = 2):
    """Measures the candidate forms on (queries, ks, efs) and keeps the
    fastest within recall_slack of exact FP32.  `forms` replaces the first
    stage's list; `refine` is how many of its best forms stage 2 varies."""
    queries = np.ascontiguousarray(queries, np.float32)
    ks = np.ascontiguousarray(ks, np.int32)
    efs = np.ascontiguousarray(efs, np.int32)
    key = (queries.shape[1], _ef_bucket(efs), int(ks.max()))
    if key in _cache and apply:
        best = _cache[key]
        _apply(index, best["spec"])
        return best
    stage1 = [(tuple(f) + (0, 0, 0))[:7] for f in (forms or STAGE1)]
    kmax = int(ks.max())
    gt_ids, gt_d, _ = index.bruteForce(queries, min(128, max(kmax, 10)))
    index.setQueryTimeParams({"algoType": "v1merge"})

    def measure(form):
        _apply(index, form)
        ids, _, _ = index.knnQueryBatchUneven(queries, ks, efs)  # warm: conversions, code objects
        el = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            ids, _, _ = index.knnQueryBatchUneven(queries, ks, efs)
            el = min(el, time.perf_counter() - t)
        launch = index.lastLaunch()
        return {"form": form[0], "spec": form, "storage": form[1], "rerank": form[2],
                "beam": form[3], "waves": form[4], "gather": form[5],
                "screen": form[6] if len(form) > 6 else 0, "seconds": el,
                "qps": len(queries) / el, "recall": _recall(ids, gt_ids, gt_d, ks),
                "kernel": launch.get("kernel"), "mode": launch.get("mode")}

    report = [measure(f) for f in stage1]
    exact = next((r for r in report if r["beam"] == 0 and r["storage"] == "fp32"), report[0])
    floor = exact["recall"] - recall_slack
    ok = sorted((r for r in report if r["recall"] >= floor), key=lambda r: -r["qps"])
    seen = {r["form"] for r in report}
    for r in ok[:refine]:
        for v in _variants(r["spec"]):
            if v[0] not in seen:
                seen.add(v[0])
                report.append(measure(v))
    ok = [r for r in report if r["recall"] >= floor]
    best = dict(max(ok, key=lambda r: r["qps"]))
    best["candidates"] = report
    best["key"] = key
    if apply:
        _apply(index, best["spec"])
        _cache[key] = best
    else:
        _apply(index, ("fp32-exact", "fp32", 0, 0, 0, 0, 0))
    return best

"""Benchmark of the HNSW query path on one MI355X.

Workload (the 1,000,000 x 128 l2 case): base vectors with integer coordinates
0..255 from a clipped half-normal (sigma 40, ~30% zeros, seed 2), an index
built by the reference nmslib with M=16, efConstruction=500 (its graph ships
as a pack under workloads/, the vectors are regenerated from the seed and
the reference-format index file is rewritten byte-for-byte, then loaded with
loadIndex), 10,000 queries from the same distribution (seed 3), k=10,
exact brute-force ground truth.

A step is one knnQueryBatch over the 10,000 queries with queries and
outputs resident in HBM.  The operating point is the smallest efSearch of a
ladder whose recall@10 (tie-aware, as the reference's eval_metrics.h counts
it) reaches 0.95.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "halfnormal1m"
PACK = os.path.join(ROOT, "workloads", "halfnormal1m_M16_efc500.pack.npz")
GT = os.path.join(ROOT, "workloads", "halfnormal1m_gt.npz")
EF_LADDER = [2560, 2816, 3072, 3328, 3584, 3840, 4096, 4352, 4608, 4736, 4864, 4992, 5120,
             5632, 6144, 7168, 8192, 10240, 12288]
SWEEP = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def recall_at_k(ids, gt_ids, gt_d, k=10):
    """Tie-aware recall@k (eval_metrics.h:112-127, eval_results.h:122-128)."""
    kth = gt_d[:, k - 1:k]
    tie = gt_d <= kth  # ground truth is exact (integer distances)
    tot = 0
    for a, g, t in zip(ids[:, :k], gt_ids, tie):
        good = set(g[t].tolist())
        tot += sum(1 for x in a.tolist() if x in good)
    return tot / (k * len(ids))


def _canon(ids, dists):
    order = np.lexsort((ids, dists), axis=-1)
    return np.take_along_axis(ids, order, -1), np.take_along_axis(dists, order, -1)


def prepare_index(tmpdir):
    from nmslib_amd import datasets as D
    from nmslib_amd import indexpack

    t = time.time()
    pk = indexpack.load_pack(PACK)
    x = D.base(WORKLOAD)
    path = os.path.join(tmpdir, "halfnormal1m.hnsw")
    indexpack.unpack(pk, x, path)
    log(f"[bench] index file written ({os.path.getsize(path) / 1e6:.0f} MB) in {time.time() - t:.1f}s")
    return path, x


def cpu_baseline(path, q, gids, gd, ef, algo, gpu_ids, gpu_dists):
    """The reference (oracle/_ref build, when present) or the oracle port timed
    on the host cores over a bounded sample of the same queries."""
    import ctypes as C

    threads = len(os.sched_getaffinity(0))
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libnmsref.so")
    kind = "reference" if os.path.exists(ref_so) else "port"
    if kind == "reference":
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        from refdrv import RefIndex
        r = RefIndex.load("l2", path, q.shape[1])
        r.set_query_params(f"efSearch={ef},algoType={algo}")

        def run(qs):
            return r.knn_query_batch(qs, 10, num_threads=threads)
    else:
        from oracle.oracle import OracleIndex
        o = OracleIndex(path)

        def run(qs):
            return o.search(qs, 10, ef, algo, num_threads=threads)

    # size the sample for ~15 s of CPU work
    t = time.time()
    run(q[:threads * 4])
    per_q = (time.time() - t) / (threads * 4)
    ns = int(min(len(q), max(threads * 8, 15.0 / max(per_q, 1e-6))))
    ns -= ns % threads or 0
    ns = max(ns, threads)
    t = time.time()
    ids, dists, counts = run(q[:ns])
    el = time.time() - t
    # exactly tied distances come back in Object-address order from the
    # reference (knnqueue.h:73-74): compare rows with ties ordered by id
    same_rows = -1
    if gpu_ids is not None:
        same_rows = int(sum(np.array_equal(a, b) and np.array_equal(c, d) for a, c, b, d in
                            zip(*_canon(ids, dists), *_canon(gpu_ids[:ns], gpu_dists[:ns]))))
    rec = recall_at_k(ids, gids[:ns], gd[:ns])
    del run
    return {"value": round(ns / el, 2), "unit": "queries/s", "cores": threads, "kind": kind,
            "sample": f"first {ns} of the {len(q)} queries, efSearch={ef}, algoType={algo}, "
                      f"k=10, {threads} threads; recall@10={rec:.4f}; rows identical to the "
                      f"GPU's (ties ordered by id): {same_rows}/{ns}"}


def _measured_traffic(path, form, ef):
    """HBM bytes per launch of `form` at `ef`, from the committed summary of
    the FETCH_SIZE / WRITE_SIZE counter runs (a line
    'traffic <form> ef=<ef> bytes=<n>'), or None when not measured."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(rf"^traffic {re.escape(form)} ef={ef} bytes=(\d+)", text, re.M)
    return int(m.group(1)) if m else None


def serve_bench(idx, q, forms, rates, duration_s, lanes=8, max_queries=1024, ef_fixed=0,
                max_ef=800, size_range=(1, 256), ks=(1, 10, 100)):
    """The serving loop over the same resident index: Poisson arrivals, each
    request 1..256 queries with k in {1, 10, 100} and ef in 50..800 (or one
    fixed ef and k = 10), admitted into free lanes as they arrive (continuous
    batching).  Per form: the sustained rate and per-request p50 / p99 at
    each offered rate, and the highest rate sustained with p99 <= 5 ms."""
    from nmslib_amd import serving

    out = []
    for name, storage, rerank, beam, waves, gather, screen in forms:
        idx.setPrecision(storage, rerank)
        idx.setBeam(beam)
        idx.setBeamWaves(waves)
        idx.setBeamGather(gather)
        idx.setScreen(screen)
        idx.setQueryTimeParams({"algoType": "v1merge"})
        srv = serving.Server(idx, lanes=lanes, max_queries=max_queries, max_ef=max_ef,
                             max_k=max(ks), use_graph=True)
        rows = []
        for rate in rates:
            ef_range = (ef_fixed, ef_fixed) if ef_fixed else (50, 800)
            reqs = serving.make_requests(len(q), rate, duration_s, seed=int(rate),
                                         size_range=size_range, ks=ks, ef_range=ef_range)
            st = serving.run_poisson(srv, q, reqs)
            st.update({"offered_qps": rate})
            rows.append({k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()})
            log(f"[bench] serve {name} ef={'50..800' if not ef_fixed else ef_fixed} "
                f"offered={rate} -> {rows[-1]}")
        srv.close()
        # sustained: served at the rate the requests arrived (the Poisson sample
        # of the offered rate) with p99 <= 5 ms
        ok = [r for r in rows if r["latency_ms_p99"] <= 5.0 and r["qps"] >= 0.97 * r["arrival_qps"]]
        out.append({"form": name, "efSearch": f"{ef_fixed}" if ef_fixed else "50..800",
                    "k": "10" if ef_fixed else "1,10,100", "lanes": lanes,
                    "max_queries_per_batch": max_queries,
                    "max_sustained_qps_at_p99_le_5ms": max((r["qps"] for r in ok), default=None),
                    "peak_qps": max(r["qps"] for r in rows), "rates": rows})
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)
    idx.setBeamWaves(0)
    idx.setBeamGather(0)
    idx.setScreen(0)
    return out


def single_query_latency(idx, q, forms, ef, n=32):
    """Device time of one query searched alone (a batch of one) at `ef`:
    the floor of any served request's latency at that operating point."""
    out = {}
    idx.setQueryTimeParams({"efSearch": ef, "algoType": "v1merge"})
    for name, storage, rerank, beam, waves, gather, screen in forms:
        idx.setPrecision(storage, rerank)
        idx.setBeam(beam)
        idx.setBeamWaves(waves)
        idx.setBeamGather(gather)
        idx.setScreen(screen)
        idx.knnQueryBatchArrays(q[:1], 10)
        ms = []
        for i in range(n):
            idx.knnQueryBatchArrays(q[i:i + 1], 10)
            ms.append(idx.lastStats()["kernel_ms"])
        out[name] = {"efSearch": ef, "queries": n, "p50_ms": round(float(np.percentile(ms, 50)), 3),
                     "p99_ms": round(float(np.percentile(ms, 99)), 3),
                     "min_ms": round(float(np.min(ms)), 3)}
        log(f"[bench] single-query latency {name}: {out[name]}")
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)
    idx.setBeamWaves(0)
    idx.setBeamGather(0)
    idx.setScreen(0)
    return out


FORMS = {
    # name: (storage, rerank multiple of k, beam width, waves per query: 0 =
    # by LDS footprint, 8 = the cooperative kernel)
    "fp32-exact": ("fp32", 0, 0, 0),
    "fp32-beam4": ("fp32", 0, 4, 0),
    "bf16-beam4-rerank": ("bf16", 2, 4, 0),
    "fp32-beam8": ("fp32", 0, 8, 0),
    "fp32-beam16": ("fp32", 0, 16, 0),
    "bf16-beam16-rerank": ("bf16", 2, 16, 0),
    "fp8-beam16-rerank": ("fp8", 2, 16, 0),
    "fp32-coop16": ("fp32", 0, 16, 8),
    "fp32-coop32": ("fp32", 0, 32, 8),
    "fp32-coop40": ("fp32", 0, 40, 8),
    "fp32-coop48": ("fp32", 0, 48, 8),
    "fp32-coop64": ("fp32", 0, 64, 8),
    "bf16-coop16-rerank": ("bf16", 2, 16, 8),
    "bf16-coop32-rerank": ("bf16", 2, 32, 8),
    "bf16-coop48-rerank": ("bf16", 2, 48, 8),
    "bf16-coop64-rerank": ("bf16", 2, 64, 8),
    # FP32 with the BF16 screen (results identical to the form without it)
    "fp32-coop48-screen": ("fp32", 0, 48, 8, 1),
    "fp32-coop64-screen": ("fp32", 0, 64, 8, 1),
    "fp32-coop64-screen-verify": ("fp32", 0, 64, 8, 2),
}


def form_spec(form):
    """(storage, rerank multiple, beam width, waves per query, screen)."""
    return (tuple(FORMS[form]) + (0,))[:5]


def measure(idx, q, gids, gd, form, args, dev, log_prefix):
    """Ladder to recall@10 >= 0.95, then a counters run and the timed steps."""
    import torch

    storage, rerank, beam, waves, screen = form_spec(form)
    idx.setPrecision(storage, rerank)
    idx.setBeam(beam)
    idx.setBeamWaves(waves)
    idx.setScreen(screen)
    nq, k = len(q), 10
    dq = torch.from_numpy(q).to(dev)
    dids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    dd = torch.empty((nq, k), dtype=torch.float32, device=dev)
    dc = torch.empty((nq,), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        idx.knnQueryBatchDevice(dq.data_ptr(), nq, k, dids.data_ptr(), dd.data_ptr(),
                                dc.data_ptr(), stream.cuda_stream)

    def run_at(ef, reps=1):
        # wall time of one batch, best of `reps` (benchmark/eval.py:485-510)
        idx.setQueryTimeParams({"efSearch": ef, "algoType": args.algo})
        el = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.time()
            step()
            torch.cuda.synchronize()
            el = min(el, time.time() - t0)
        return recall_at_k(dids.cpu().numpy(), gids, gd), el

    curve = []
    if args.sweep:
        for ef in SWEEP:
            r, el = run_at(ef, reps=3)
            curve.append({"ef": ef, "recall": round(r, 4), "qps": round(nq / el, 1)})
            log(f"{log_prefix} sweep ef={ef} recall={r:.4f} qps={nq / el:.0f}")
    ef = args.ef
    if not ef:
        for cand in EF_LADDER:
            r, el = run_at(cand)
            log(f"{log_prefix} ladder ef={cand} recall={r:.4f} ({nq / el:.0f} q/s)")
            if r >= 0.95:
                ef = cand
                break
        else:
            ef = EF_LADDER[-1]
    idx.setQueryTimeParams({"efSearch": ef, "algoType": args.algo})

    # counters + per-query device latency (untimed run)
    idx.collectStats(True)
    step()
    torch.cuda.synchronize()
    st = idx.lastStats()
    launch = idx.lastLaunch()
    lat = idx.lastQueryTimesUs(nq)
    phases = idx.lastPhaseUs() if args.algo != "old" else None
    idx.collectStats(False)
    ids_np = dids.cpu().numpy()
    dists_np = dd.cpu().numpy()
    recall = recall_at_k(ids_np, gids, gd)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.time()
    for _ in range(args.steps):
        step()
        torch.cuda.synchronize()
        kernel_ms.append(idx.lastStats()["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = time.time() - t0
    # the same batch through host buffers (knnQueryBatch semantics: queries
    # copied in, results copied out over PCIe); reported beside, never `value`
    t1 = time.time()
    idx.knnQueryBatchArrays(q, k)
    host_qps = nq / (time.time() - t1)

    # algorithmic bytes of one launch: adjacency ints read, every distance
    # evaluation's vector, the rerank's FP32 vectors, queries in, results out;
    # under the BF16 screen a screened candidate reads its BF16 vector, and
    # its FP32 one only if it passed
    dim = q.shape[1]
    vec_bytes = {"fp32": 4, "bf16": 2, "fp8": 1}[storage] * dim + (4 if storage == "fp8" else 0)
    rerank_bytes = nq * min(ef, rerank * k) * dim * 4 if rerank else 0
    n_vec = st["dist_level0"] + st["dist_upper"] - st["screened"] + st["screen_pass"]
    alg_bytes = (st["list_ints_read"] * 4 + n_vec * vec_bytes + st["screened"] * 2 * dim
                 + rerank_bytes + nq * dim * 4 + nq * (k * 8 + 4))
    avg_kernel_s = float(np.mean(kernel_ms)) / 1000
    res = {
        "form": form, "efSearch": ef, "recall_at_10": round(recall, 4),
        "qps": round(nq * args.steps / elapsed, 1),
        "qps_host_buffers": round(host_qps, 1),
        "ms_per_step": round(elapsed * 1000 / args.steps, 3),
        "avg_kernel_ms": round(avg_kernel_s * 1000, 3),
        "latency_us_p50": round(float(np.percentile(lat, 50)), 1),
        "latency_us_p99": round(float(np.percentile(lat, 99)), 1),
        "per_query": {"expansions": round(st["expansions"] / nq, 1),
                      "dist_evals": round((st["dist_level0"] + st["dist_upper"]) / nq, 1),
                      "tie_path_expansions": round(st["tie_paths"] / nq, 1),
                      "bf16_screened": round(st["screened"] / nq, 1),
                      "screen_passed_to_fp32": round(st["screen_pass"] / nq, 1)},
        "alg_bytes_per_launch": int(alg_bytes),
        "achieved_gbs": round(alg_bytes / avg_kernel_s / 1e9, 1),
        "launch": launch,
        "curve": curve,
    }
    if phases:
        tot = sum(phases.values()) or 1.0
        res["step_phase_share"] = {p: round(v / tot, 3) for p, v in phases.items()}
    log(f"{log_prefix} {form}: {res}")
    idx.setScreen(0)
    return res, ids_np, dists_np


def gmm_workload(args):
    """The reduced-precision negdotprod configuration at full size: 10,000,000 x
    96 vectors of the normalised 64-component Gaussian mixture, an index built
    here by createIndex (M=24, efConstruction=200, all host cores), 10,000
    queries, k=10.  FP32 exact V1Merge against BF16 and FP8-e4m3 traversal with
    an exact FP32 rerank of all ef candidates: recall@10 against the device
    brute force, and the distances' error against FP64 on a query sample.
    Prints one JSON line."""
    import nmslib_amd as nmslib
    from nmslib_amd import datasets as D

    name = "gmm10m"
    n = args.gmm_n or D.WORKLOADS[name][1]
    t = time.time()
    x = D.base(name, n)
    q = D.queries(name)
    log(f"[gmm] {n:,} x {x.shape[1]} generated in {time.time() - t:.0f}s")
    threads = len(os.sched_getaffinity(0))
    idx = nmslib.init(method="hnsw", space="negdotprod")
    idx.addDataPointBatch(x)
    t = time.time()
    idx.createIndex({"M": 24, "efConstruction": 200, "indexThreadQty": threads},
                    print_progress=True)
    build_s = time.time() - t
    log(f"[gmm] createIndex {build_s:.0f}s on {threads} threads")
    k = 10
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)
    gt, gd, _ = idx.bruteForce(q, 100)
    bf_ms = idx.lastStats()["kernel_ms"]
    sample = np.arange(0, len(q), 625)  # 16 queries
    x64 = x.astype(np.float64)

    def fp64(rows, ids):
        out = np.zeros(ids.shape, np.float64)
        for i, r in enumerate(rows):
            ok = ids[i] >= 0
            out[i, ok] = -(x64[ids[i, ok]] @ q[r].astype(np.float64))
        return out

    def recall(ids):
        kth = gd[:, k - 1:k]
        hit = 0
        for a, g, gdr, th in zip(ids[:, :k], gt, gd, kth):
            good = set(g[gdr <= th[0] + 1e-6 * max(1.0, abs(th[0]))].tolist())
            hit += sum(1 for v in a.tolist() if v in good)
        return hit / (k * len(ids))

    rows = []
    for storage, beam, ef in [("fp32", 0, 100), ("fp32", 0, 400), ("fp32", 8, 800),
                              ("bf16", 0, 100), ("bf16", 0, 400), ("bf16", 8, 800),
                              ("fp8", 0, 100), ("fp8", 0, 400), ("fp8", 8, 800)]:
        mult = ef // k if storage != "fp32" else 0  # rerank all ef candidates
        idx.setBeam(beam)
        idx.setQueryTimeParams({"efSearch": ef, "algoType": "v1merge"})
        row = {"storage": storage, "beam": beam, "efSearch": ef, "rerank_candidates": mult * k}
        if storage != "fp32":
            # the traversal's own distances (no rerank) against FP64
            idx.setPrecision(storage, 0)
            li, ld, _ = idx.knnQueryBatchArrays(q[sample], k)
            err = np.abs(ld.astype(np.float64) - fp64(sample, li))
            row["traversal_err_mean"] = float(f"{err.mean():.3e}")
            row["traversal_err_max"] = float(f"{err.max():.3e}")
        idx.setPrecision(storage, mult)
        idx.knnQueryBatchArrays(q, k)  # warm-up
        best = float("inf")
        for _ in range(3):
            ids, dists, counts = idx.knnQueryBatchArrays(q, k)
            best = min(best, idx.lastStats()["kernel_ms"])
        err = np.abs(dists[sample].astype(np.float64) - fp64(sample, ids[sample]))
        row.update({"qps": round(len(q) / best * 1e3, 1), "kernel_ms": round(best, 3),
                    "recall_at_10": round(recall(ids), 4),
                    "returned_err_max": float(f"{err.max():.3e}")})
        log(f"[gmm] {row}")
        rows.append(row)
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    best_lowp = max((r for r in rows if r["storage"] != "fp32" and r["recall_at_10"] >= 0.95),
                    key=lambda r: r["qps"], default=None)
    out = {"metric": "QPS at recall@10 (10M x 96 negdotprod, BF16/FP8 traversal + FP32 rerank)",
           "value": best_lowp["qps"] if best_lowp else None, "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic (seeded 64-component Gaussian mixture)",
           "dtype": best_lowp["storage"] if best_lowp else None,
           "config": {"workload": f"{n:,} x 96 negdotprod GMM, M=24 efC=200, 10,000 queries, k=10",
                      "build_s": round(build_s, 1), "build_threads": threads,
                      "bruteforce_ms": round(bf_ms, 2), "error_sample_queries": len(sample),
                      "rows": rows}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ef", type=int, default=0, help="fixed efSearch (0: smallest of the ladder "
                    "reaching recall@10 >= 0.95)")
    ap.add_argument("--algo", default="v1merge")
    ap.add_argument("--forms", default="fp32-exact,fp32-beam16,bf16-beam16-rerank,fp32-coop64,"
                                       "fp32-coop64-screen,bf16-coop64-rerank",
                    help=f"comma list of {sorted(FORMS)}; the headline is the fastest")
    ap.add_argument("--nq", type=int, default=10000)
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the efSearch 16..4096 recall/QPS curve of each form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", default="5000,20000,50000,100000,200000,400000,800000",
                    help="offered query rates (q/s) of the serving run, '' to skip")
    ap.add_argument("--serve-hi", default="10000,20000,50000,80000",
                    help="offered rates of the serving run at the headline efSearch")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-forms", default="",
                    help="comma list of serving form names to run (default: all)")
    ap.add_argument("--serve-lanes", type=int, default=8,
                    help="concurrent batch lanes of the serving run")
    ap.add_argument("--serve-batch", type=int, default=1024,
                    help="most queries one lane launches at once")
    ap.add_argument("--build-threads", type=int, default=0,
                    help="build the index with this library's createIndex on N host threads "
                         "(M=16, efConstruction=500, as the reference-built pack) instead of "
                         "loading the reference's file; 0 = load the reference's file")
    ap.add_argument("--visited-marks", type=int, default=0,
                    help="visited marks: 0 = by size, 1 = epoch bytes, 2 = bits (setVisitedMarks)")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="only a 1 GiB copy and one search launch (for rocprofv3 --pmc runs)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "hnsw_search_kernel_halfnormal1m_traffic.md"),
                    help="HBM bytes per launch measured from FETCH_SIZE/WRITE_SIZE runs")
    ap.add_argument("--workload", default=WORKLOAD, choices=[WORKLOAD, "gmm10m"],
                    help="halfnormal1m: the 1M x 128 l2 headline (default); gmm10m: the "
                         "10M x 96 negdotprod reduced-precision case, index built here")
    ap.add_argument("--gmm-n", type=int, default=0, help="gmm10m: base vectors (0 = 10,000,000)")
    args = ap.parse_args()
    if args.workload == "gmm10m" and args.gpus == 1:
        gmm_workload(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the query path "
              f"would shard by replicating the index per GPU (one process each)", file=sys.stderr)
        sys.exit(2)

    import torch

    import nmslib_amd as nmslib
    from nmslib_amd import datasets as D

    tmpdir = tempfile.mkdtemp(prefix="nmsamd_bench_")
    q = D.queries(WORKLOAD, args.nq)
    gt = np.load(GT)
    gids, gd = gt["ids"][:args.nq], gt["dists"][:args.nq]

    idx = nmslib.init(method="hnsw", space="l2")
    build_s = None
    if args.build_threads > 0:
        # the graph built here on the host (createIndex), saved for the CPU leg
        idx.addDataPointBatch(D.base(WORKLOAD))
        t = time.time()
        idx.createIndex({"M": 16, "efConstruction": 500, "indexThreadQty": args.build_threads},
                        print_progress=True)
        build_s = time.time() - t
        log(f"[bench] createIndex {build_s:.1f}s on {args.build_threads} threads {idx.info()}")
        path = os.path.join(tmpdir, "halfnormal1m_built.hnsw")
        idx.saveIndex(path)
    else:
        path, _ = prepare_index(tmpdir)
        t = time.time()
        idx.loadIndex(path)
        log(f"[bench] loadIndex {time.time() - t:.1f}s {idx.info()}")
    dev = torch.device("cuda:0")
    idx.setVisitedMarks(args.visited_marks)

    if args.pmc_pass:
        # short fixed pass for hardware-counter runs (rocprofv3 --pmc): a plain
        # copy of known size (the counters' scale), then one search launch
        nbytes = 1 << 30
        src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        form = args.forms.split(",")[-1]
        storage, rerank, beam, waves, screen = form_spec(form)
        idx.setPrecision(storage, rerank)
        idx.setBeam(beam)
        idx.setBeamWaves(waves)
        idx.setScreen(screen)
        idx.setQueryTimeParams({"efSearch": args.ef or 4736, "algoType": args.algo})
        ids, dists, counts = idx.knnQueryBatchArrays(q, 10)
        log(f"[bench] pmc pass: copy {nbytes} B, one {form} launch, recall "
            f"{recall_at_k(ids, gids, gd):.4f}")
        return

    forms = [f for f in args.forms.split(",") if f]
    results, exact_out = {}, None
    for form in forms:
        res, ids_np, dists_np = measure(idx, q, gids, gd, form, args, dev, "[bench]")
        results[form] = res
        if form == "fp32-exact":
            exact_out = (res["efSearch"], ids_np, dists_np)
    # the headline is the fastest form computing in FP32 (the reference's
    # arithmetic); reduced-precision forms are reported beside it
    fp32 = [r for f, r in results.items() if FORMS[f][0] == "fp32"]
    head = max(fp32 or results.values(), key=lambda r: r["qps"])
    lowp = {f: {"qps": r["qps"], "recall_at_10": r["recall_at_10"],
                "ratio_to_headline": round(r["qps"] / head["qps"], 3)}
            for f, r in results.items() if FORMS[f][0] != "fp32"}

    # the dense contraction of the path: exact kNN by brute force on MFMA
    # (device seq_search), timed by HIP events; FP32 MFMA peak 157.3 TFLOP/s
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)
    bf_ids, bf_d, _ = idx.bruteForce(q, 10)
    bf_ms = idx.lastStats()["kernel_ms"]
    bf_flops = 2.0 * len(q) * 1_000_000 * 128
    ci, cd = _canon(bf_ids, bf_d)
    bf_exact = bool(np.array_equal(cd, gd[:, :10]) and np.array_equal(ci, gids[:, :10]))
    brute = {"kernel": "bf_tile_kernel + bf_merge_kernel + bf_rescore_kernel", "queries": len(q),
             "k": 10, "rescanned_queries": idx.lastBruteForceRescans(),
             "ms": round(bf_ms, 3), "tflops": round(bf_flops / bf_ms / 1e9, 1),
             "frac_of_fp32_mfma_peak": round(bf_flops / bf_ms / 1e9 / 157.3, 4),
             "equals_ground_truth": bf_exact}
    log(f"[bench] brute force {brute}")

    serve = None
    if args.serve or args.serve_hi:
        rates = [float(r) for r in args.serve.split(",") if r]
        # (name, storage, rerank multiple, beam width, waves per query, gather:
        # 2 keeps the cooperative kernel at one block per CU, BF16 screen)
        serve_forms = [("fp32 exact V1Merge", "fp32", 0, 0, 0, 0, 0),
                       ("fp32 beam16", "fp32", 0, 16, 0, 0, 0),
                       ("bf16 beam16 + fp32 rerank x2", "bf16", 2, 16, 0, 0, 0),
                       ("fp32 coop64", "fp32", 0, 64, 8, 0, 0),
                       ("fp32 coop64 + bf16 screen", "fp32", 0, 64, 8, 0, 1),
                       ("bf16 coop64 + fp32 rerank x2", "bf16", 2, 64, 8, 0, 0),
                       ("bf16 coop64 one block per CU + fp32 rerank x2", "bf16", 2, 64, 8, 2, 0)]
        if args.serve_forms:
            serve_forms = [f for f in serve_forms if f[0] in args.serve_forms.split(",")]
        mix = serve_bench(idx, q, serve_forms, rates, args.serve_seconds, args.serve_lanes,
                          args.serve_batch) if rates else []
        # the recall@10 >= 0.95 operating point: every request at the headline
        # efSearch, k = 10
        hi_rates = [float(r) for r in args.serve_hi.split(",") if r]
        hi = serve_bench(idx, q, [f for f in serve_forms if f[3] > 0], hi_rates, args.serve_seconds,
                         args.serve_lanes,
                         args.serve_batch, ef_fixed=head["efSearch"], max_ef=head["efSearch"],
                         ks=(10,)) if hi_rates else []
        alone = single_query_latency(idx, q, serve_forms, head["efSearch"])
        serve = {"poisson_mix": mix, "at_recall_0_95": hi, "single_query_latency": alone}

    cpu = None
    if not args.no_cpu:
        if exact_out is not None:
            ef_cpu, ids_cmp, d_cmp = exact_out
        else:
            ef_cpu, ids_cmp, d_cmp = head["efSearch"], None, None
        cpu = cpu_baseline(path, q, gids, gd, ef_cpu, args.algo, ids_cmp, d_cmp)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": "QPS at recall@10>=0.95 (1M x 128 L2, 1 GPU) + p50/p99 latency",
        "value": head["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16", "fp8": "fp8"}[FORMS[head["form"]][0]],
        "data": ("synthetic (seeded generator; index built by createIndex in "
                 f"{build_s:.0f} s on {args.build_threads} host threads)" if build_s is not None
                 else "synthetic (seeded generator; reference-built index)"),
        "config": {"workload": "1,000,000 x 128 l2 half-normal u8-valued, M=16 efC=500, "
                               "10,000 queries, k=10",
                   "headline_form": head["form"],
                   "headline_note": ("FP32 distances and results, identical to the form without "
                                     "the screen; each candidate's BF16 copy gives a proven lower "
                                     "bound on its FP32 distance, and candidates whose bound "
                                     "reaches the admission key skip their FP32 vector"
                                     if form_spec(head["form"])[4] else
                                     "FP32 distances and results"),
                   "efSearch": head["efSearch"],
                   "algoType": args.algo, "recall_at_10": head["recall_at_10"],
                   "latency_us_p50": head["latency_us_p50"],
                   "latency_us_p99": head["latency_us_p99"],
                   "latency_kind": "per-query device search time inside the batch",
                   "forms": results,
                   "reduced_precision_forms": lowp},
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": _measured_traffic(args.traffic, head["form"], head["efSearch"]),
                     "kernel": head["launch"]["kernel"],
                     "avg_kernel_ms": head["avg_kernel_ms"],
                     "alg_bytes_per_launch": head["alg_bytes_per_launch"]},
        "cpu_baseline": cpu,
        "serving": serve,
        "bruteforce": brute,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

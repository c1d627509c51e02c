"""Writes the regular-format fixtures reg_<name>.npz of tests/golden/.

Runs only where the reference sources exist (``make -C oracle ref``).  Each
case runs in a child process of its own, so the reference's level generator
starts from seed 0: the reference builds the index single-threaded with
skip_optimized_index=1, which keeps the regular (non-optimized) form
(hnsw.cc:356-360), saves it (SaveRegularIndexBin, hnsw.cc:810-841; the file
holds no vectors) and answers knnQueryBatch on it with
baseSearchAlgorithmV1Merge / baseSearchAlgorithmOld (hnsw.cc:1076-1300),
which use the space's own distance with the data point as left operand.

Each fixture holds the saved index file's bytes, the queries and, for each
(efSearch, k, algoType), the reference's ids / distances / counts.  The data
points are regenerated from the named seeded generator.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# name: (space, generator, n, dim, seed, nq, qseed, index params)
CASES = {
    "r_u128_l2": ("l2", "uniform", 2000, 128, 51, 100, 52, "M=16,efConstruction=100"),
    "r_u30_l2": ("l2", "uniform", 1200, 30, 53, 100, 54, "M=8,efConstruction=60"),
    "r_hn36_l2": ("l2", "halfnormal_u8", 1500, 36, 55, 100, 56, "M=8,efConstruction=60"),
    "r_dec100_cos": ("cosinesimil", "decaying_normal", 2000, 100, 57, 100, 58,
                     "M=16,efConstruction=100"),
    "r_gmm96_nd": ("negdotprod", "gmm", 2000, 96, 59, 100, 60, "M=12,efConstruction=100"),
    "r_u67_l1": ("l1", "uniform", 1500, 67, 61, 100, 62, "M=8,efConstruction=60"),
    "r_u20_linf": ("linf", "uniform", 1000, 20, 63, 100, 64, "M=6,efConstruction=40"),
}

# (efSearch, k, algoType)
RUNS = [(10, 10, "hybrid"), (32, 10, "hybrid"), (100, 10, "hybrid"), (16, 1, "hybrid"),
        (200, 50, "v1merge"), (64, 10, "old"), (1000, 10, "hybrid")]


def one(name):
    """Builds, saves and queries one case (in a fresh process)."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, HERE)
    from nmslib_amd import datasets as D
    from refdrv import RefIndex
    space, g, n, dim, seed, nq, qseed, params = CASES[name]
    x = getattr(D, g)(n, dim, seed)
    q = getattr(D, g)(nq, dim, qseed)
    if space == "cosinesimil":
        q[0] = 0.0  # an all-zero query: the space's cosine reports distance 1
    ref = RefIndex.build(space, x, params + ",skip_optimized_index=1,indexThreadQty=1")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "index.hnsw")
        ref.save(path)
        raw = np.fromfile(path, dtype=np.uint8)
    assert raw[:4].view(np.uint32)[0] == 0, "expected the regular format"
    out = {"index": raw, "queries": q}
    for ef, k, algo in RUNS:
        ref.set_query_params(f"efSearch={ef},algoType={algo}")
        ids, dists, counts = ref.knn_query_batch(q, k, num_threads=1)
        tag = f"ef{ef}_k{k}_{algo}"
        out[f"{tag}_ids"], out[f"{tag}_dists"], out[f"{tag}_counts"] = ids, dists, counts
    np.savez_compressed(os.path.join(HERE, f"reg_{name}.npz"), **out)


def main(only=None):
    meta = {}
    for name, (space, g, n, dim, seed, nq, qseed, params) in CASES.items():
        if only and name not in only:
            continue
        subprocess.run([sys.executable, "-c",
                        f"import sys; sys.path.insert(0, {HERE!r}); "
                        f"import make_regular_goldens as m; m.one({name!r})"], check=True)
        meta[name] = dict(space=space, generator=g, n=n, dim=dim, seed=seed, nq=nq, qseed=qseed,
                          params=params, runs=[list(r) for r in RUNS])
        print(name, "done", flush=True)
    mpath = os.path.join(HERE, "regular_cases.json")
    old = json.load(open(mpath)) if os.path.exists(mpath) else {}
    old.update(meta)
    with open(mpath, "w") as f:
        json.dump(old, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)

"""Writes the full-size workloads under workloads/ (git-ignored; they travel
with the tree to the GPU host, where the benchmark and the full-size tests
read them).

    python tests/golden/make_workloads.py halfnormal1m [decay2m ...]

For each workload the base vectors come from the seeded generator
(nmslib_amd/datasets.py), the index is built by the reference nmslib through
oracle/_ref/ref_build (``make -C oracle ref``; oracle/ref_build_main.cc drives
addDataPointBatch / createIndex / saveIndex as python_bindings/nmslib.cc does)
and kept as a pack (nmslib_amd/indexpack.py: the file without its vectors,
which unpack back byte-identical).  The exact ground truth is an FP64 brute
force over all base vectors, ranked by (distance, id) as gold_standard.h:73-76
orders it.

  halfnormal1m   1,000,000 x 128 l2, M=16, efConstruction=500, 10,000 queries
                 -> halfnormal1m_M16_efc500.pack.npz, halfnormal1m_gt.npz
                    (ids int32 / squared-L2 dists float32 of the 100 nearest;
                    the coordinates are integers, so every distance is exact)
  decay2m        2,000,000 x 100 cosinesimil, M=32, efConstruction=400
                 -> decay2m_M32_efc400.pack.npz
  gmm2m          2,000,000 x 96 negdotprod, M=24, efConstruction=200
                 -> gmm2m_M24_efc200.pack.npz
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nmslib_amd import datasets as D  # noqa: E402
from nmslib_amd import indexpack  # noqa: E402

OUT = os.path.join(ROOT, "workloads")
REF_BUILD = os.path.join(ROOT, "oracle", "_ref", "ref_build")

# name: (space, index params, pack file)
SPECS = {
    "halfnormal1m": ("l2", "M=16,efConstruction=500", "halfnormal1m_M16_efc500.pack.npz"),
    "decay2m": ("cosinesimil", "M=32,efConstruction=400", "decay2m_M32_efc400.pack.npz"),
    "gmm2m": ("negdotprod", "M=24,efConstruction=200", "gmm2m_M24_efc200.pack.npz"),
}


def log(*a):
    print(*a, flush=True)


def build_pack(name: str, threads: int) -> None:
    space, params, fname = SPECS[name]
    x = D.base(name)
    with tempfile.TemporaryDirectory(prefix="nmsamd_wl_") as td:
        raw = os.path.join(td, "data.f32")
        x.tofile(raw)
        path = os.path.join(td, "index.hnsw")
        t = time.time()
        subprocess.run([REF_BUILD, space, raw, str(x.shape[0]), str(x.shape[1]),
                        f"{params},indexThreadQty={threads}", path], check=True)
        log(f"[{name}] reference build {time.time() - t:.0f} s on {threads} threads")
        pk = indexpack.pack(path)
        out = os.path.join(OUT, fname)
        indexpack.save_pack(pk, out)
        back = os.path.join(td, "back.hnsw")
        indexpack.unpack(indexpack.load_pack(out), x, back)
        same = subprocess.run(["cmp", "-s", path, back]).returncode == 0
        log(f"[{name}] pack {os.path.getsize(out) / 1e6:.0f} MB, round trip byte-identical: {same}")
        if not same:
            raise SystemExit("pack round trip differs")


def l2_ground_truth(x: np.ndarray, q: np.ndarray, k: int = 100, chunk: int = 256):
    """FP64 squared-L2 brute force, (distance, id) order."""
    x64 = x.astype(np.float64)
    xn = (x64 * x64).sum(1)
    ids = np.empty((len(q), k), np.int32)
    dists = np.empty((len(q), k), np.float32)
    for s in range(0, len(q), chunk):
        qq = q[s:s + chunk].astype(np.float64)
        d = xn[None, :] - 2.0 * (qq @ x64.T) + (qq * qq).sum(1)[:, None]
        part = np.argpartition(d, k, axis=1)[:, :k + 1]
        for r in range(len(qq)):
            # every candidate tied with the k-th must be seen to order by id
            kth = np.sort(d[r, part[r]])[k - 1]
            cand = np.nonzero(d[r] <= kth)[0]
            o = np.lexsort((cand, d[r, cand]))[:k]
            ids[s + r] = cand[o]
            dists[s + r] = d[r, cand[o]]
        log(f"[gt] {min(s + chunk, len(q))}/{len(q)}")
    return ids, dists


def main(names, threads):
    os.makedirs(OUT, exist_ok=True)
    for name in names:
        build_pack(name, threads)
        if name == "halfnormal1m":
            x = D.base(name)
            q = D.queries(name)
            ids, dists = l2_ground_truth(x, q)
            # integer coordinates: the FP64 values are exact integers
            assert np.all(dists == np.rint(dists))
            np.savez_compressed(os.path.join(OUT, "halfnormal1m_gt.npz"), ids=ids, dists=dists)
            log("[halfnormal1m] ground truth written")


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--threads=")]
    thr = [int(a.split("=")[1]) for a in sys.argv[1:] if a.startswith("--threads=")]
    main(args or ["halfnormal1m"], thr[0] if thr else os.cpu_count())

"""Writes the index-construction fixtures bld_<name>.npz of tests/golden/.

Runs only where the reference sources exist: ``make -C oracle ref`` compiles
them into oracle/_ref/ref_build (oracle/ref_build_main.cc), which builds one
index with the reference's Hnsw::CreateIndex and saves it.  Every case runs
in a process of its own, single-threaded (indexThreadQty=1), so the
reference's per-thread level generator starts from seed 0 each time.

Each fixture is the saved index as a pack (nmslib_amd/indexpack.py; vectors
regenerated from the named seeded generator) with the unused slots of the
upper-level lists set to zero: the reference leaves those bytes
uninitialised (hnsw.cc:455-465 mallocs the lists without clearing them).

The cases cover what the fixtures idx_<name>.npz (written by make_goldens.py)
do not: the other neighbour-selection rules (delaunay_type 0, 1 and 3), both
post-processing modes, L2 with a dimension that is not a multiple of 16
(L2SqrExt), custom M / maxM / maxM0 / mult, and integer data whose distances
tie often, 1 to 3 points, and data where every point appears four times.
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
sys.path.insert(0, ROOT)

from nmslib_amd import datasets as D  # noqa: E402
from nmslib_amd import indexpack  # noqa: E402

REF_BUILD = os.path.join(ROOT, "oracle", "_ref", "ref_build")

# name: (space, generator, n, dim, seed, index params)
CASES = {
    "b_u20_l2": ("l2", "uniform", 1000, 20, 4, "M=6,efConstruction=40"),
    "b_hn16_l2": ("l2", "halfnormal_u8", 1200, 16, 9, "M=8,efConstruction=64"),
    "b_d0_l2": ("l2", "uniform", 800, 16, 10, "M=8,efConstruction=40,delaunay_type=0"),
    "b_d1_l2": ("l2", "uniform", 800, 16, 11, "M=8,efConstruction=40,delaunay_type=1"),
    "b_post2_l2": ("l2", "uniform", 800, 16, 12, "M=8,efConstruction=40,post=2"),
    "b_post1_l2": ("l2", "uniform", 800, 16, 13, "M=8,efConstruction=40,post=1"),
    "b_custom_cos": ("cosinesimil", "decaying_normal", 1000, 24, 5,
                     "M=10,maxM=7,maxM0=15,mult=0.6,efConstruction=60"),
    "b_d1_nd": ("negdotprod", "gmm", 1000, 16, 6, "M=12,efConstruction=60,delaunay_type=1"),
    "b_n1_l2": ("l2", "uniform", 1, 8, 71, "M=4,efConstruction=10"),
    "b_n2_l2": ("l2", "uniform", 2, 8, 72, "M=4,efConstruction=10"),
    "b_n3_cos": ("cosinesimil", "uniform", 3, 8, 73, "M=4,efConstruction=10"),
    # heuristic 3 on tie-free data (the reference's hash-set order only
    # decides among equal distances)
    "b_d3_l2": ("l2", "uniform", 800, 16, 75, "M=8,efConstruction=40,delaunay_type=3"),
    "b_d3_post2_l2": ("l2", "uniform", 600, 16, 76, "M=8,efConstruction=40,delaunay_type=3,post=2"),
    "b_d3_cos": ("cosinesimil", "decaying_normal", 800, 24, 77, "M=10,efConstruction=50,delaunay_type=3"),
    # every point four times over: all-zero distances between copies
    "b_dup4_l2": ("l2", "uniform", 200, 8, 74, "M=6,efConstruction=20", 4),
}


def points(case):
    """The case's data points: the generator's rows, each repeated `repeat` times."""
    space, g, n, dim, seed, params = case[:6]
    rep = case[6] if len(case) > 6 else 1
    return np.repeat(getattr(D, g)(n // rep, dim, seed), rep, axis=0)


def zero_unused_upper(pk):
    per = int(pk["hdr"][6]) + 1  # maxM + 1
    up = pk["upper"].reshape(-1, per).copy()
    for row in up:
        row[1 + row[0]:] = 0
    pk["upper"] = up.reshape(-1)
    return pk


def main(only=None):
    meta = {}
    for name, case in CASES.items():
        space, g, n, dim, seed, params = case[:6]
        if only and name not in only:
            continue
        x = points(case)
        with tempfile.TemporaryDirectory() as td:
            raw = os.path.join(td, "x.f32")
            x.astype(np.float32).tofile(raw)
            path = os.path.join(td, "index.hnsw")
            subprocess.run([REF_BUILD, space, raw, str(n), str(dim),
                            params + ",indexThreadQty=1", path], check=True)
            pk = zero_unused_upper(indexpack.pack(path))
        indexpack.save_pack(pk, os.path.join(HERE, f"bld_{name}.npz"))
        meta[name] = dict(space=space, generator=g, n=n, dim=dim, seed=seed, params=params,
                          repeat=case[6] if len(case) > 6 else 1)
        print(name, "done", flush=True)
    mpath = os.path.join(HERE, "build_cases.json")
    old = json.load(open(mpath)) if os.path.exists(mpath) else {}
    old.update(meta)
    with open(mpath, "w") as f:
        json.dump(old, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)

"""Loads the committed golden fixtures of tests/golden/ (written from the
reference by tests/golden/make_goldens.py) and rebuilds their index files."""
from __future__ import annotations

import functools
import json
import os
import tempfile

import numpy as np

from nmslib_amd import datasets as D
from nmslib_amd import indexpack

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "fixtures.json")))
NAMES = sorted(META)
_TMP = tempfile.mkdtemp(prefix="nmsamd_fx_")


def base_vectors(name):
    m = META[name]
    return getattr(D, m["generator"])(m["n"], m["dim"], m["seed"])


@functools.lru_cache(maxsize=None)
def index_path(name):
    """Reference-format index file of fixture `name` (byte-identical to the
    file the reference saved)."""
    path = os.path.join(_TMP, f"{name}.hnsw")
    if not os.path.exists(path):
        pk = indexpack.load_pack(os.path.join(GOLDEN, f"idx_{name}.npz"))
        indexpack.unpack(pk, base_vectors(name), path)
    return path


@functools.lru_cache(maxsize=None)
def results(name):
    with np.load(os.path.join(GOLDEN, f"res_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def runs(name):
    return [tuple(r) for r in META[name]["runs"]]


def tag(ef, k, algo):
    return f"ef{ef}_k{k}_{algo}"


def recall_tie_aware(ids, gt_ids, gt_dists, k, rel=1e-6):
    """Recall@k as the reference counts it (eval_metrics.h:112-127,
    eval_results.h:122-128): an answer counts when it is among the exact k
    or its exact distance ties the k-th exact distance."""
    tot = 0.0
    for a, g, gd in zip(ids, gt_ids, gt_dists):
        kth = gd[k - 1]
        ok = set(int(x) for x in g[:k])
        ties = set(int(x) for x, d in zip(g, gd) if abs(d - kth) <= rel * max(1.0, abs(kth)))
        good = ok | ties
        tot += sum(1 for x in a[:k] if int(x) in good) / k
    return tot / len(ids)


def canon(ids, dists):
    """Rows re-ordered by (distance, id).  The reference returns exactly tied
    distances in the order of their Object heap addresses (knnqueue.h:73-74,
    objects created one by one in LoadOptimizedIndex, hnsw.cc:1055-1071); that
    order follows the allocator's state, so parity compares tied runs as sets."""
    ids = np.asarray(ids)
    dists = np.asarray(dists)
    order = np.lexsort((ids, dists), axis=-1)
    return np.take_along_axis(ids, order, -1), np.take_along_axis(dists, order, -1)

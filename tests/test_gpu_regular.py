"""Device search of regular (non-optimized) indices against the reference's
answers (tests/golden/reg_<name>.npz): baseSearchAlgorithmV1Merge and
baseSearchAlgorithmOld (hnsw.cc:1076-1300) with the space's own distance (the
square root for l2, the cosine of the raw vectors, -dot), the data point as
left operand.  Bar: ids and FP32 distances bit-exact, exactly tied distances
compared as sets (the reference orders them by heap address)."""
import os
import tempfile
import threading

import numpy as np
import pytest

import nmslib_amd as nm
from tests import fixtures as F
from tests import regular_fixtures as R
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

_TMP = tempfile.mkdtemp(prefix="nmsamd_greg_")


def _check_runs(idx, name):
    fx = R.fixture(name)
    q = fx["queries"]
    for ef, k, algo in R.runs(name):
        idx.setQueryTimeParams({"efSearch": ef, "algoType": algo})
        ids, dists, counts = idx.knnQueryBatchArrays(q, k)
        t = F.tag(ef, k, algo)
        print(f"{name} {t}", flush=True)
        np.testing.assert_array_equal(counts, fx[t + "_counts"], err_msg=t)
        ci, cd = F.canon(ids, dists)
        ri, rd = F.canon(fx[t + "_ids"], fx[t + "_dists"])
        np.testing.assert_array_equal(ci, ri, err_msg=t)
        np.testing.assert_array_equal(cd.view(np.uint32), rd.view(np.uint32), err_msg=t)


@pytest.mark.parametrize("name", R.NAMES)
def test_regular_index_matches_reference(name):
    require_gpu()
    m = R.META[name]
    path = os.path.join(_TMP, f"{name}.reg")
    R.write_index(name, path)
    idx = nm.init(method="hnsw", space=m["space"])
    idx.addDataPointBatch(R.base_vectors(name))
    idx.loadIndex(path)
    _check_runs(idx, name)


@pytest.mark.parametrize("name", ["r_u128_l2", "r_dec100_cos", "r_u67_l1"])
def test_regular_build_searches_like_reference(name):
    require_gpu()
    m = R.META[name]
    holder = {}

    def run():  # a fresh thread: its level generator starts from seed 0
        try:
            idx = nm.init(method="hnsw", space=m["space"])
            idx.addDataPointBatch(R.base_vectors(name))
            idx.createIndex(m["params"].split(",") + ["skip_optimized_index=1", "indexThreadQty=1"])
            holder["idx"] = idx
        except Exception as e:
            holder["err"] = e

    t = threading.Thread(target=run)
    t.start()
    t.join()
    if "err" in holder:
        raise holder["err"]
    _check_runs(holder["idx"], name)


def test_regular_index_rejects_reduced_precision():
    require_gpu()
    name = "r_u30_l2"
    path = os.path.join(_TMP, "lowp.reg")
    R.write_index(name, path)
    idx = nm.init(method="hnsw", space="l2")
    idx.addDataPointBatch(R.base_vectors(name))
    idx.loadIndex(path)
    idx.setPrecision("bf16", 2)
    with pytest.raises(ValueError, match="needs an optimized index"):
        idx.knnQueryBatchArrays(R.fixture(name)["queries"], 10)

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


@pytest.mark.parametrize("beam", [0, 4])
def test_regular_index_serves_like_the_batch_entry_point(beam):
    """The serving lanes over a regular index: a request's answers equal the
    plain batch entry point's (exact mode, and beam mode, which uses the same
    kernels with the space's distance)."""
    require_gpu()
    from nmslib_amd import serving
    name = "r_u128_l2"
    path = os.path.join(_TMP, "serve.reg")
    R.write_index(name, path)
    idx = nm.init(method="hnsw", space="l2")
    idx.addDataPointBatch(R.base_vectors(name))
    idx.loadIndex(path)
    idx.setBeam(beam)
    q = R.fixture(name)["queries"][:40]
    ks = np.full(len(q), 10, np.int32)
    efs = np.full(len(q), 64, np.int32)
    srv = serving.Server(idx, lanes=1, max_queries=64, max_ef=64, max_k=10, use_graph=False)
    srv.submit(0, q, ks, efs)
    si, sd, sc = srv.fetch(0)
    srv.close()
    ids, dists, counts = idx.knnQueryBatchUneven(q, ks, efs)
    np.testing.assert_array_equal(sc, counts)
    np.testing.assert_array_equal(si, ids)
    np.testing.assert_array_equal(sd.view(np.uint32), dists.view(np.uint32))
    if beam == 0:  # exact mode: the same rows as the plain batch call
        ci, cd = F.canon(ids[:, :10], dists[:, :10])
        idx.setQueryTimeParams({"efSearch": 64, "algoType": "v1merge"})
        ri, rd, rc = idx.knnQueryBatchArrays(q, 10)
        np.testing.assert_array_equal(ci, F.canon(ri, rd)[0])
    idx.setPrecision("bf16", 2)
    with pytest.raises(ValueError, match="needs an optimized index"):
        serving.Server(idx, lanes=1, max_queries=64, max_ef=64, max_k=10, use_graph=False)

"""The serving loop over the 1,000,000 x 128 l2 benchmark index (the
reference-built pack under workloads/): Poisson request streams of 1-256
queries with k in {1, 10, 100} and efSearch in 50..800, served by lanes that
run concurrently on the device; every sampled request's rows equal the same
request searched alone through the batch entry point, at two offered rates
and for the exact FP32 and the BF16 beam + FP32 rerank forms."""
import os
import tempfile

import numpy as np
import pytest

from nmslib_amd import serving
from tests.gpu_util import load

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACK = os.path.join(ROOT, "workloads", "halfnormal1m_M16_efc500.pack.npz")


@pytest.fixture(scope="module")
def big():
    if not os.path.exists(PACK):
        pytest.skip("1M workload pack not present")
    from nmslib_amd import datasets as D
    from nmslib_amd import indexpack
    path = os.path.join(tempfile.mkdtemp(prefix="nmsamd_srv_"), "halfnormal1m.hnsw")
    indexpack.unpack(indexpack.load_pack(PACK), D.base("halfnormal1m"), path)
    idx = load("halfnormal1m", path, "l2")
    idx.setQueryTimeParams({"algoType": "v1merge"})
    yield idx, D.queries("halfnormal1m")
    os.remove(path)


@pytest.mark.parametrize("rate", [20_000, 100_000])
@pytest.mark.parametrize("storage,rerank,beam", [("fp32", 0, 0), ("bf16", 2, 16)])
def test_fullsize_server_mix_equals_alone(big, rate, storage, rerank, beam):
    idx, pool = big
    idx.setPrecision(storage, rerank)
    idx.setBeam(beam)
    srv = serving.Server(idx, lanes=8, max_queries=1024, max_ef=800, max_k=100, use_graph=True)
    reqs = serving.make_requests(len(pool), rate_qps=rate, duration_s=0.5, seed=rate + beam)
    st = serving.run_poisson(srv, pool, reqs, keep_results=True)
    srv.close()
    print(f"{storage} beam {beam} offered {rate} q/s: {st}", flush=True)
    assert st["requests"] == len(reqs)
    for r in reqs[::max(1, len(reqs) // 24)]:
        q = pool[r.rows]
        ids, dists, counts = idx.knnQueryBatchUneven(q, np.full(len(q), r.k, np.int32),
                                                     np.full(len(q), r.ef, np.int32))
        gi, gd, gc = r.result
        np.testing.assert_array_equal(gc, counts, err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gi, ids[:, :r.k], err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gd, dists[:, :r.k], err_msg=f"request {r.rid}")
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)


def test_fullsize_server_two_wave_lanes_at_the_recall_point(big):
    # every request at efSearch 4736 (recall@10 >= 0.95): the lanes take the
    # two-wave kernel, as the batch entry point does at that footprint
    idx, pool = big
    idx.setPrecision("bf16", 2)
    idx.setBeam(16)
    srv = serving.Server(idx, lanes=4, max_queries=512, max_ef=4736, max_k=10, use_graph=True)
    assert srv.info()["beam_kernel"] == 2
    reqs = serving.make_requests(len(pool), rate_qps=20_000, duration_s=0.3, seed=11,
                                 size_range=(1, 128), ks=(10,), ef_range=(4736, 4736))
    st = serving.run_poisson(srv, pool, reqs, keep_results=True)
    srv.close()
    print(f"bf16 beam 16, ef 4736, two-wave lanes, offered 20k q/s: {st}", flush=True)
    for r in reqs[::max(1, len(reqs) // 12)]:
        q = pool[r.rows]
        ids, dists, counts = idx.knnQueryBatchUneven(q, np.full(len(q), r.k, np.int32),
                                                     np.full(len(q), r.ef, np.int32))
        gi, gd, gc = r.result
        np.testing.assert_array_equal(gc, counts, err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gi, ids[:, :r.k], err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gd, dists[:, :r.k], err_msg=f"request {r.rid}")
    idx.setPrecision("fp32", 0)
    idx.setBeam(0)


@pytest.mark.parametrize("storage,rerank,gather,screen", [("fp32", 0, 0, 0), ("fp32", 0, 0, 1),
                                                          ("bf16", 2, 0, 0), ("bf16", 2, 2, 0)])
def test_fullsize_server_coop_lanes_at_the_recall_point(big, storage, rerank, gather, screen):
    # every request at efSearch 4736 on the cooperative kernel (beam 64, a
    # block of eight waves per query; two blocks per CU, or one with gather 2;
    # FP32 also with the BF16 screen): the served rows equal the batch entry
    # point's for the same requests
    idx, pool = big
    idx.setPrecision(storage, rerank)
    idx.setBeam(64)
    idx.setBeamWaves(8)
    idx.setBeamGather(gather)
    idx.setScreen(screen)
    srv = serving.Server(idx, lanes=4, max_queries=256, max_ef=4736, max_k=10, use_graph=True)
    assert srv.info()["beam_kernel"] == (3 if gather == 2 else 4)
    reqs = serving.make_requests(len(pool), rate_qps=20_000, duration_s=0.3, seed=13,
                                 size_range=(1, 64), ks=(10,), ef_range=(4736, 4736))
    st = serving.run_poisson(srv, pool, reqs, keep_results=True)
    srv.close()
    print(f"{storage} coop 64, ef 4736, offered 20k q/s: {st}", flush=True)
    for r in reqs[::max(1, len(reqs) // 12)]:
        q = pool[r.rows]
        ids, dists, counts = idx.knnQueryBatchUneven(q, np.full(len(q), r.k, np.int32),
                                                     np.full(len(q), r.ef, np.int32))
        gi, gd, gc = r.result
        np.testing.assert_array_equal(gc, counts, err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gi, ids[:, :r.k], err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(gd, dists[:, :r.k], err_msg=f"request {r.rid}")
    idx.setPrecision("fp32", 0)
    idx.setScreen(0)
    idx.setBeamGather(0)
    idx.setBeamWaves(0)
    idx.setBeam(0)

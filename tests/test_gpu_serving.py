"""The continuous-batching server: results are the same alone and in the
mix, with and without hipGraph replay, under a Poisson request stream."""
import numpy as np
import pytest

from nmslib_amd import serving
from tests import fixtures as F
from tests.gpu_util import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("beam", [0, 4])
def test_server_mix_equals_alone(use_graph, beam):
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    idx.setBeam(beam)
    pool = F.results(name)["queries"]
    srv = serving.Server(idx, lanes=3, max_queries=512, max_ef=400, max_k=100,
                         use_graph=use_graph)
    info = srv.info()
    assert info["lanes"] == 3 and info["graph"] == int(use_graph)
    reqs = serving.make_requests(len(pool), rate_qps=40000, duration_s=0.25, seed=7,
                                 size_range=(1, 64), ks=(1, 10, 100), ef_range=(16, 400))
    assert len(reqs) > 50
    stats = serving.run_poisson(srv, pool, reqs, keep_results=True)
    assert stats["requests"] == len(reqs)
    print(stats)
    # each request alone, through the plain batch entry point
    for r in reqs[::7]:
        q = pool[r.rows]
        ids, dists, counts = idx.knnQueryBatchUneven(q, np.full(len(q), r.k, np.int32),
                                                     np.full(len(q), r.ef, np.int32))
        gi, gd, gc = r.result
        np.testing.assert_array_equal(gc, counts)
        np.testing.assert_array_equal(gi, ids[:, :r.k])
        np.testing.assert_array_equal(gd, dists[:, :r.k])
    srv.close()


def test_server_rejects_out_of_budget_batches():
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    srv = serving.Server(idx, lanes=1, max_queries=8, max_ef=64, max_k=10, use_graph=False)
    q = F.results(name)["queries"][:9]
    with pytest.raises(ValueError):
        srv.submit(0, q, np.full(9, 10, np.int32), np.full(9, 32, np.int32))
    with pytest.raises(ValueError):
        srv.submit(0, q[:2], np.full(2, 10, np.int32), np.full(2, 65, np.int32))
    srv.submit(0, q[:2], np.full(2, 10, np.int32), np.full(2, 64, np.int32))
    with pytest.raises(ValueError):  # lane busy
        srv.submit(0, q[:2], np.full(2, 10, np.int32), np.full(2, 64, np.int32))
    ids, dists, counts = srv.fetch(0)
    assert counts.tolist() == [10, 10]
    srv.close()


def test_autotune_picks_a_form_within_the_recall_slack():
    from nmslib_amd import autotune
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    q = F.results(name)["queries"]
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
    efs = rng.integers(32, 800, size=len(q)).astype(np.int32)
    best = autotune.autotune(idx, q, ks, efs, reps=1)
    exact = next(c for c in best["candidates"] if c["form"] == "fp32-exact")
    assert best["recall"] >= exact["recall"] - 0.003
    assert best["qps"] >= max(c["qps"] for c in best["candidates"]
                              if c["recall"] >= exact["recall"] - 0.003) - 1e-9
    print({c["form"]: (round(c["qps"]), round(c["recall"], 4)) for c in best["candidates"]})
    # the choice is applied: a second call hits the cache
    again = autotune.autotune(idx, q, ks, efs)
    assert again["form"] == best["form"]


def test_index_cannot_be_replaced_under_an_open_server():
    # a server holds the index's device pointers (in captured graphs too):
    # loading or building another index on the handle waits for its close
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    srv = serving.Server(idx, lanes=1, max_queries=8, max_ef=64, max_k=10, use_graph=True)
    with pytest.raises(ValueError, match="close the servers"):
        idx.loadIndex(F.index_path(name))
    idx.addDataPointBatch(F.base_vectors(name)[:100])
    with pytest.raises(ValueError, match="close the servers"):
        idx.createIndex({"M": 8})
    srv.close()
    idx.loadIndex(F.index_path(name))  # allowed again
    assert idx.info()["n"] == F.META[name]["n"]


def test_server_keeps_its_form_when_the_index_setters_change():
    # a server searches with the precision, rerank and beam it was opened
    # with (its LDS layout and captured graphs follow them); changing them on
    # the index afterwards affects the batch entry points only
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    pool = F.results(name)["queries"][:64]
    ks = np.full(len(pool), 10, np.int32)
    efs = np.full(len(pool), 100, np.int32)
    alone = idx.knnQueryBatchUneven(pool, ks, efs)
    for use_graph in (False, True):
        srv = serving.Server(idx, lanes=1, max_queries=64, max_ef=128, max_k=10,
                             use_graph=use_graph)
        idx.setPrecision("bf16", 2)
        idx.setBeam(16)
        srv.submit(0, pool, ks, efs)
        ids, dists, counts = srv.fetch(0)
        idx.setPrecision("fp32", 0)
        idx.setBeam(0)
        srv.close()
        np.testing.assert_array_equal(counts, alone[2])
        np.testing.assert_array_equal(ids, alone[0][:, :10])
        np.testing.assert_array_equal(dists, alone[1][:, :10])


def test_bf16_server_opened_before_any_batch_search():
    # the server is the first user of the BF16 copy: its conversion finishes
    # before the lanes' first batch reads it
    name = "hn128_l2"  # integer coordinates: BF16 traversal equals FP32
    idx = load(name, F.index_path(name), "l2")
    idx.setPrecision("bf16", 0)
    srv = serving.Server(idx, lanes=2, max_queries=200, max_ef=100, max_k=10, use_graph=True)
    pool = F.results(name)["queries"]
    ks = np.full(len(pool), 10, np.int32)
    efs = np.full(len(pool), 100, np.int32)
    srv.submit(1, pool, ks, efs)
    ids, dists, counts = srv.fetch(1)
    srv.close()
    idx.setPrecision("fp32", 0)
    ref = idx.knnQueryBatchUneven(pool, ks, efs)
    np.testing.assert_array_equal(counts, ref[2])
    np.testing.assert_array_equal(ids, ref[0][:, :10])
    np.testing.assert_array_equal(dists, ref[1][:, :10])

"""Parity at the full 1,000,000 x 128 size (the reference-built M=16,
efC=500 index of the benchmark, shipped as a pack under workloads/): the
device search equals the oracle row for row on a query sample, the device
brute force equals the committed exact ground truth, and the beam / BF16
forms keep the recall of the exact search."""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import load

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACK = os.path.join(ROOT, "workloads", "halfnormal1m_M16_efc500.pack.npz")
GT = os.path.join(ROOT, "workloads", "halfnormal1m_gt.npz")


@pytest.fixture(scope="module")
def big():
    if not (os.path.exists(PACK) and os.path.exists(GT)):
        pytest.skip("1M workload pack not present")
    from nmslib_amd import datasets as D
    from nmslib_amd import indexpack
    path = os.path.join(tempfile.mkdtemp(prefix="nmsamd_full_"), "halfnormal1m.hnsw")
    indexpack.unpack(indexpack.load_pack(PACK), D.base("halfnormal1m"), path)
    q = D.queries("halfnormal1m", 200)
    gt = np.load(GT)
    idx = load("halfnormal1m", path, "l2")
    return path, idx, q, gt["ids"][:200], gt["dists"][:200]


def _canon(ids, d):
    o = np.lexsort((ids, d), axis=-1)
    return np.take_along_axis(ids, o, -1), np.take_along_axis(d, o, -1)


def _recall(ids, gids, gd, k=10):
    kth = gd[:, k - 1:k]
    tot = 0
    for a, g, t in zip(ids[:, :k], gids, gd <= kth):
        good = set(g[t].tolist())
        tot += sum(1 for x in a.tolist() if x in good)
    return tot / (k * len(ids))


@pytest.mark.parametrize("ef", [64, 512])
def test_fullsize_exact_search_equals_oracle(big, ef):
    path, idx, q, _, _ = big
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    idx.setQueryTimeParams({"efSearch": ef})
    ids, dists, counts = idx.knnQueryBatchArrays(q, 10)
    oi, od, oc = O.OracleIndex(path).search(q, 10, ef, num_threads=8)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(counts, oc)


def test_fullsize_exact_search_at_the_recall_point_equals_oracle(big):
    # efSearch 4736, where recall@10 first reaches 0.95: the exact search's
    # rows, distances and counts equal the oracle's on a sample of 48 queries
    path, idx, q, _, _ = big
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    idx.setQueryTimeParams({"efSearch": 4736, "algoType": "v1merge"})
    ids, dists, counts = idx.knnQueryBatchArrays(q[:48], 10)
    oi, od, oc = O.OracleIndex(path).search(q[:48], 10, 4736, "v1merge", num_threads=8)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(counts, oc)


@pytest.mark.parametrize("beam,waves", [(4, 1), (16, 1), (4, 2), (16, 2), (16, 8), (32, 8),
                                        (64, 8)])
def test_fullsize_beam_search_equals_model(big, beam, waves):
    # at the benchmark's efSearch the array spans 74 chunks of 64 and a step
    # merges up to a few hundred items (beam 64: about a thousand): the device
    # equals the CPU model of the beam search row for row, with one wave per
    # query, two, and the cooperative block of eight
    path, idx, q, _, _ = big
    idx.setBeam(beam)
    idx.setBeamWaves(waves)
    idx.setPrecision("fp32", 0)
    idx.setQueryTimeParams({"efSearch": 4736, "algoType": "v1merge"})
    ids, dists, counts = idx.knnQueryBatchArrays(q[:48], 10)
    print(f"beam {beam} waves {waves}: {idx.lastLaunch()}", flush=True)
    idx.setBeam(0)
    idx.setBeamWaves(0)
    oi, od, oc = O.OracleIndex(path).search_beam(q[:48], 10, 4736, beam, num_threads=8)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(counts, oc)


@pytest.mark.parametrize("beam,screen", [(64, 2), (48, 1)])
def test_fullsize_screened_beam_search_equals_model(big, beam, screen):
    # the BF16 screen of the cooperative FP32 search leaves out candidates
    # whose proven lower bound reaches the admission key: the rows equal the
    # FP32 beam model's; screen 2 also checks every bound against the FP32
    # distance on the device (a failure raises)
    path, idx, q, _, _ = big
    idx.setBeam(beam)
    idx.setBeamWaves(8)
    idx.setScreen(screen)
    idx.setPrecision("fp32", 0)
    idx.setQueryTimeParams({"efSearch": 4736, "algoType": "v1merge"})
    idx.collectStats(True)
    try:
        ids, dists, counts = idx.knnQueryBatchArrays(q[:48], 10)
        st = idx.lastStats()
    finally:
        idx.collectStats(False)
        idx.setScreen(0)
        idx.setBeam(0)
        idx.setBeamWaves(0)
    print(f"beam {beam} screen {screen}: screened {st['screened']} passed {st['screen_pass']} "
          f"of {st['dist_level0']}", flush=True)
    assert st["screened"] > st["dist_level0"] // 2 and st["screen_pass"] < st["screened"]
    oi, od, oc = O.OracleIndex(path).search_beam(q[:48], 10, 4736, beam, num_threads=8)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(counts, oc)


def test_fullsize_bruteforce_equals_ground_truth(big):
    _, idx, q, gids, gd = big
    ids, dists, counts = idx.bruteForce(q[:64], 100)
    ci, cd = _canon(ids, dists)
    np.testing.assert_array_equal(cd, gd[:64])
    np.testing.assert_array_equal(ci, gids[:64])


def test_fullsize_fast_forms_keep_recall(big):
    _, idx, q, gids, gd = big
    idx.setQueryTimeParams({"efSearch": 1024, "algoType": "v1merge"})
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    base = _recall(idx.knnQueryBatchArrays(q, 10)[0], gids, gd)
    idx.setBeam(4)
    idx.setPrecision("bf16", 2)
    fast = _recall(idx.knnQueryBatchArrays(q, 10)[0], gids, gd)
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    print(f"recall@10 at ef 1024: exact {base:.4f}, bf16 beam4 rerank {fast:.4f}")
    assert fast >= base - 0.003

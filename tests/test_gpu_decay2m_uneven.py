"""The 2,000,000 x 100 cosinesimil case at full size: an index built by the
reference nmslib (M=32, efConstruction=400; its graph ships as a pack under
workloads/, the vectors are regenerated from their seed), one uneven batch of
8,192 queries with k drawn from {1, 10, 100} and efSearch from 32..800.

The device's exact mode equals the oracle row for row on a sample of the
batch, beam mode equals the beam model on the same sample, and the recall of
both against the device brute force is reported.  Skipped when the pack is
absent (it is too large to keep in the tree)."""
import os
import tempfile
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import load

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PACK = os.path.join(ROOT, "workloads", "decay2m_M32_efc400.pack.npz")
SAMPLE = slice(0, 8192, 128)  # 64 queries of the batch


@pytest.fixture(scope="module")
def decay2m():
    if not os.path.exists(PACK):
        pytest.skip("2M cosine workload pack not present")
    from nmslib_amd import datasets as D
    from nmslib_amd import indexpack
    path = os.path.join(tempfile.mkdtemp(prefix="nmsamd_decay2m_"), "decay2m.hnsw")
    indexpack.unpack(indexpack.load_pack(PACK), D.base("decay2m"), path)
    q = D.queries("decay2m")
    rng = np.random.default_rng(8)
    ks = rng.choice(np.array([1, 10, 100], np.int32), size=len(q)).astype(np.int32)
    efs = rng.integers(32, 801, size=len(q)).astype(np.int32)
    idx = load("decay2m", path, "cosinesimil")
    gt, gd, _ = idx.bruteForce(q, 100)
    yield path, idx, q, ks, efs, gt
    os.remove(path)


def _recall(ids, ks, gt):
    out = {}
    for kk in (1, 10, 100):
        sel = np.nonzero(ks == kk)[0]
        hit = sum(len(set(ids[i, :kk].tolist()) & set(gt[i, :kk].tolist())) for i in sel)
        out[kk] = hit / (kk * len(sel))
    return out


def _timed_uneven(idx, q, ks, efs):
    idx.knnQueryBatchUneven(q, ks, efs)  # warm-up (first use builds reduced-precision copies)
    t = time.time()
    ids, dists, counts = idx.knnQueryBatchUneven(q, ks, efs)
    wall = time.time() - t
    return ids, dists, counts, idx.lastStats()["kernel_ms"], wall


def test_decay2m_uneven_exact_equals_oracle(decay2m):
    path, idx, q, ks, efs, gt = decay2m
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    idx.setQueryTimeParams({"algoType": "v1merge"})
    ids, dists, counts, ms, wall = _timed_uneven(idx, q, ks, efs)
    rec = _recall(ids, ks, gt)
    print(f"decay2m exact uneven: {len(q) / ms * 1e3:.0f} q/s (kernel {ms:.1f} ms, wall {wall * 1e3:.1f} ms),"
          f" recall by k {rec}", flush=True)
    oracle = O.OracleIndex(path)
    for i in range(len(q))[SAMPLE]:
        k, ef = int(ks[i]), int(efs[i])
        oi, od, oc = oracle.search(q[i:i + 1], k, ef, "v1merge")
        assert counts[i] == oc[0], i
        np.testing.assert_array_equal(ids[i, :k], oi[0], err_msg=f"query {i} k={k} ef={ef}")
        np.testing.assert_array_equal(dists[i, :k].view(np.uint32), od[0].view(np.uint32))


@pytest.mark.parametrize("storage,rerank", [("fp32", 0), ("bf16", 2)])
def test_decay2m_uneven_beam16(decay2m, storage, rerank):
    path, idx, q, ks, efs, gt = decay2m
    idx.setBeam(16)
    idx.setPrecision(storage, rerank)
    idx.setQueryTimeParams({"algoType": "v1merge"})
    ids, dists, counts, ms, wall = _timed_uneven(idx, q, ks, efs)
    rec = _recall(ids, ks, gt)
    launch = idx.lastLaunch()
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    print(f"decay2m {storage} beam16 rerank x{rerank} uneven: {len(q) / ms * 1e3:.0f} q/s "
          f"(kernel {ms:.1f} ms, wall {wall * 1e3:.1f} ms), recall by k {rec}, {launch}", flush=True)
    if storage != "fp32":
        return
    oracle = O.OracleIndex(path)
    w = max(1, min(16, 512 // oracle.maxM0))
    for i in range(len(q))[SAMPLE]:
        k, ef = int(ks[i]), int(efs[i])
        oi, od, oc = oracle.search_beam(q[i:i + 1], k, ef, w)
        assert counts[i] == oc[0], i
        np.testing.assert_array_equal(ids[i, :k], oi[0], err_msg=f"query {i} k={k} ef={ef}")
        np.testing.assert_array_equal(dists[i, :k].view(np.uint32), od[0].view(np.uint32))


def test_decay2m_autotuner_pick_keeps_pace_with_the_hand_set_form(decay2m):
    # the autotuner (storage x beam width, then waves per query, gather width
    # and rerank width for the best forms) calibrated on a quarter of the
    # batch: on the whole batch its pick keeps the exact search's recall
    # (within the slack) and is at least as fast as the hand-set BF16 beam 8
    # + rerank x2 form, one of the forms it weighs
    from nmslib_amd import autotune
    path, idx, q, ks, efs, gt = decay2m
    cal = slice(0, len(q), 4)
    best = autotune.autotune(idx, q[cal], ks[cal], efs[cal], reps=2)
    print({c["form"]: (round(c["qps"]), round(c["recall"], 4)) for c in best["candidates"]},
          flush=True)
    pick = _timed_uneven(idx, q, ks, efs)
    autotune._apply(idx, ("bf16-beam8-rerank2", "bf16", 2, 8, 0, 0))
    hand = _timed_uneven(idx, q, ks, efs)
    autotune._apply(idx, ("fp32-exact", "fp32", 0, 0, 0, 0))
    exact = _timed_uneven(idx, q, ks, efs)

    def mean_recall(ids):
        r = _recall(ids, ks, gt)
        return sum(r[kk] * np.sum(ks == kk) for kk in r) / len(ks)

    r_pick, r_hand, r_exact = mean_recall(pick[0]), mean_recall(hand[0]), mean_recall(exact[0])
    print(f"decay2m autotuned {best['form']}: kernel {pick[3]:.1f} ms, recall {r_pick:.4f}; "
          f"bf16 beam 8: {hand[3]:.1f} ms, {r_hand:.4f}; fp32 exact: {exact[3]:.1f} ms, "
          f"{r_exact:.4f}", flush=True)
    assert r_pick >= r_exact - 0.006
    if r_hand >= r_exact - 0.003:
        assert pick[3] <= hand[3] * 1.05



@pytest.mark.parametrize("storage,rerank", [("fp32", 0), ("bf16", 2)])
def test_decay2m_two_waves_equal_one_wave(decay2m, storage, rerank):
    # the whole uneven batch with two waves per query (768 blocks, three per
    # CU, dimension 100: the instantiations whose hand-off barrier lacked its
    # LDS wait) returns exactly the one-wave kernel's rows
    path, idx, q, ks, efs, gt = decay2m
    idx.setPrecision(storage, rerank)
    idx.setBeam(8)
    idx.setQueryTimeParams({"algoType": "v1merge"})
    idx.setBeamWaves(1)
    one = idx.knnQueryBatchUneven(q, ks, efs)
    idx.setBeamWaves(2)
    two = idx.knnQueryBatchUneven(q, ks, efs)
    launch = idx.lastLaunch()
    idx.setBeamWaves(0)
    idx.setBeam(0)
    idx.setPrecision("fp32", 0)
    assert launch["mode"] == "beam (two waves)" and launch["blocks"] > 512, launch
    for a, b in zip(one, two):
        np.testing.assert_array_equal(a, b)

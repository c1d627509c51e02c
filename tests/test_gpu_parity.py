"""Parity of the HIP search with the reference (golden fixtures written from
the compiled reference) and with the CPU oracle, through the C ABI.

Bar: ids and FP32 distances bit-exact, with exactly tied distances compared
as sets (the reference orders them by heap address, knnqueue.h:73-74)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import fixtures as F
from tests.gpu_util import load

pytestmark = pytest.mark.gpu



def assert_same(ids, dists, counts, rids, rdists, rcounts, msg):
    np.testing.assert_array_equal(counts, rcounts, err_msg=msg)
    ci, cd = F.canon(ids, dists)
    ri, rd = F.canon(rids, rdists)
    np.testing.assert_array_equal(ci, ri, err_msg=msg)
    np.testing.assert_array_equal(cd.view(np.uint32), rd.view(np.uint32), err_msg=msg)


@pytest.mark.parametrize("name", F.NAMES)
def test_gpu_matches_reference_goldens(name):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    res = F.results(name)
    q = res["queries"]
    for ef, k, algo in F.runs(name):  # V1Merge, SearchOld (old / ef >= 1000)
        idx.setQueryTimeParams({"efSearch": ef, "algoType": algo})
        ids, dists, counts = idx.knnQueryBatchArrays(q, k)
        t = F.tag(ef, k, algo)
        print(f"{name} {t}", flush=True)
        assert_same(ids, dists, counts, res[t + "_ids"], res[t + "_dists"], res[t + "_counts"],
                    f"{name} {t}")


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "dec100_cos", "gmm96_nd", "u67_l1",
                                  "u64_linf"])
@pytest.mark.parametrize("ef,k", [(24, 10), (48, 5), (200, 50), (300, 1)])
def test_gpu_matches_oracle(name, ef, k):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"], ef)
    q = F.results(name)["queries"]
    ids, dists, counts = idx.knnQueryBatchArrays(q, k)
    oi, od, oc = O.OracleIndex(F.index_path(name)).search(q, k, ef)
    np.testing.assert_array_equal(ids, oi)  # the oracle orders ties by internal index too
    np.testing.assert_array_equal(dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(counts, oc)


def test_gpu_uniform10k_full_batch():
    # 10,000 x 128 uniform, l2, M=16, efC=200; 1,000 queries, k=10, ef=100
    name = "u10k_l2"
    idx = load(name, F.index_path(name), "l2", 100)
    res = F.results(name)
    idx.collectStats(True)
    ids, dists, counts = idx.knnQueryBatchArrays(res["queries"], 10)
    assert_same(ids, dists, counts, res["ef100_k10_hybrid_ids"], res["ef100_k10_hybrid_dists"],
                res["ef100_k10_hybrid_counts"], "u10k_l2")
    st = idx.lastStats()
    _, _, _, ost = O.OracleIndex(F.index_path(name)).search(res["queries"], 10, 100,
                                                            with_stats=True)
    assert st["expansions"] == ost["expansions"]
    assert st["dist_level0"] == ost["dist_level0"]
    assert st["dist_upper"] == ost["dist_upper"]
    assert st["list_ints_read"] == ost["list_ints_read"]


def test_gpu_uneven_batch_equals_single_runs():
    name = "u128_l2"
    idx = load(name, F.index_path(name), "l2")
    q = F.results(name)["queries"]
    rng = np.random.default_rng(0)
    ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
    efs = rng.integers(16, 400, size=len(q)).astype(np.int32)
    ids, dists, counts = idx.knnQueryBatchUneven(q, ks, efs)
    oracle = O.OracleIndex(F.index_path(name))
    for i in range(len(q)):
        oi, od, oc = oracle.search(q[i:i + 1], int(ks[i]), int(efs[i]))
        assert counts[i] == oc[0]
        np.testing.assert_array_equal(ids[i, :oc[0]], oi[0, :oc[0]])
        np.testing.assert_array_equal(dists[i, :oc[0]], od[0, :oc[0]])


def test_gpu_edge_cases():
    name = "tiny_l2"
    idx = load(name, F.index_path(name), "l2", 10)
    q = F.results(name)["queries"]
    # k above ef and above the index size: the array grows to max(ef, k) but
    # only min(size, ef) items are expanded (hnsw_distfunc_opt.cc:200-211)
    ids, dists, counts = idx.knnQueryBatchArrays(q[:3], 100)
    oi, od, oc = O.OracleIndex(F.index_path(name)).search(q[:3], 100, 10)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(ids, oi)
    for row, c in zip(ids, counts):
        assert np.all(row[c:] == -1) and len(set(row[:c].tolist())) == c
    # with ef >= n every element comes back once
    idx.setQueryTimeParams({"efSearch": 100})
    ids, dists, counts = idx.knnQueryBatchArrays(q[:3], 100)
    assert np.all(counts == 60)
    for row in ids:
        assert sorted(row[:60].tolist()) == list(range(60))
    # the list form holds only the found results (nmslib.cc:170-184)
    rows = idx.knnQueryBatch(q[:3], 100)
    for (ri, rd), i, d in zip(rows, ids, dists):
        assert ri.dtype == np.int32 and rd.dtype == np.float32 and len(ri) == len(rd) == 60
        np.testing.assert_array_equal(ri, i[:60])
        np.testing.assert_array_equal(rd, d[:60])
    idx.setQueryTimeParams({"efSearch": 10})
    # k = 0 and an empty batch
    ids, dists, counts = idx.knnQueryBatchArrays(q[:3], 0)
    assert counts.tolist() == [0, 0, 0]
    assert idx.knnQueryBatch(np.zeros((0, 128), np.float32), 5) == []
    # knnQuery is a batch of one, list-of-vectors input and Fortran order agree
    one = idx.knnQuery(q[0], k=5)
    lst = idx.knnQueryBatch([list(v) for v in q[:2]], k=5)
    fort = idx.knnQueryBatch(np.asfortranarray(q[:2]), k=5)
    np.testing.assert_array_equal(one[0], lst[0][0])
    np.testing.assert_array_equal(lst[1][0], fort[1][0])
    with pytest.raises(ValueError):
        idx.knnQueryBatch(np.zeros((2, 7), np.float32), 5)


def test_gpu_zero_cosine_query_gives_distance_one():
    # bindings_test.py:73-91: an all-zero query on cosinesimil returns distance 1
    name = "dec100_cos"
    idx = load(name, F.index_path(name), "cosinesimil", 10)
    ids, dists = idx.knnQuery(np.zeros(100, np.float32), k=3)
    assert dists.tolist() == [1.0, 1.0, 1.0]


def test_gpu_save_index_round_trip(tmp_path):
    name = "gmm96_nd"
    idx = load(name, F.index_path(name), "negdotprod")
    out = tmp_path / "copy.hnsw"
    idx.saveIndex(str(out))
    assert open(out, "rb").read() == open(F.index_path(name), "rb").read()
    idx2 = load(name, str(out), "negdotprod", 64)
    idx.setQueryTimeParams({"efSearch": 64})
    q = F.results(name)["queries"][:20]
    a = idx.knnQueryBatchArrays(q, 10)
    b = idx2.knnQueryBatchArrays(q, 10)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_gpu_wrong_space_for_index_is_rejected():
    name = "gmm96_nd"
    with pytest.raises(RuntimeError, match="does not match space l2"):
        load(name, F.index_path(name), "l2")


def _exact_dist(func, q, x):
    return np.float32(O.dist(func, q, x))


@pytest.mark.parametrize("name", ["u128_l2", "dec100_cos", "gmm96_nd"])
def test_gpu_beam1_equals_exact_without_ties(name):
    # beam mode with one expansion per step is V1Merge with ties ordered by id;
    # these fixtures have no tied distances, so results must match exactly
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"], 64)
    q = F.results(name)["queries"][1:]  # skip the all-zero cosine query (all ties)
    exact = idx.knnQueryBatchArrays(q, 10)
    idx.setBeam(1)
    beam = idx.knnQueryBatchArrays(q, 10)
    for x, y in zip(exact, beam):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "dec100_cos", "gmm96_nd"])
@pytest.mark.parametrize("beam", [2, 4])
def test_gpu_beam_mode_quality(name, beam):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"], 100)
    res = F.results(name)
    q = res["queries"]
    ex_ids, ex_d, _ = idx.knnQueryBatchArrays(q, 10)
    idx.setBeam(beam)
    ids, dists, counts = idx.knnQueryBatchArrays(q, 10)
    assert np.all(counts == 10)
    x = F.base_vectors(name)
    oracle = O.OracleIndex(F.index_path(name))
    blk = oracle.level0_block().reshape(oracle.n, oracle.elem_bytes)
    vec = blk[:, 16:16 + 4 * oracle.dim].copy().view(np.float32)  # as stored (normalised for cosine)
    for i in range(len(q)):
        row = ids[i]
        assert len(set(row.tolist())) == 10
        assert np.all(np.diff(dists[i]) >= 0)
        qq = O.normalize(q[i]) if m["space"] == "cosinesimil" else q[i]
        for j in (0, 9):
            assert dists[i, j] == _exact_dist(oracle.func, qq, vec[row[j]])
    # recall against brute force stays within 0.3 points of exact mode (or above)
    gt, gd = O.brute_force(x, q, 100, m["space"])
    r_exact = F.recall_tie_aware(ex_ids, gt, gd, 10, rel=1e-5)
    r_beam = F.recall_tie_aware(ids, gt, gd, 10, rel=1e-5)
    assert r_beam >= r_exact - 0.003, (r_beam, r_exact)


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "dec100_cos", "gmm96_nd", "u10k_l2",
                                  "u67_l1", "u64_linf"])
@pytest.mark.parametrize("beam", [1, 2, 5, 8, 16])
@pytest.mark.parametrize("waves", [1, 2])
def test_gpu_beam_mode_equals_model(name, beam, waves):
    # the device's beam mode (bitonic sort + bitmap merge of each step's
    # admitted items) against the CPU model of the same algorithm: ids,
    # distances and work counters identical; array sizes from one to a few
    # chunks, k > ef, and steps of up to 512 entries; one wave per query, and
    # two (the second wave computing part of the distances and rebuilding
    # half of the array)
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    oracle = O.OracleIndex(F.index_path(name))
    w = max(1, min(beam, 512 // oracle.maxM0))
    q = F.results(name)["queries"]
    idx.setBeam(beam)
    idx.setBeamWaves(waves)
    idx.collectStats(True)
    for k, ef in [(10, 64), (100, 32), (10, 300), (1, 700)]:
        idx.setQueryTimeParams({"efSearch": ef, "algoType": "v1merge"})
        ids, dists, counts = idx.knnQueryBatchArrays(q, k)
        st = idx.lastStats()
        rids, rdists, rcounts, rst = oracle.search_beam(q, k, ef, w, num_threads=8, with_stats=True)
        msg = f"{name} beam={beam} waves={waves} k={k} ef={ef} {idx.lastLaunch()['kernel']}"
        print(msg, flush=True)
        np.testing.assert_array_equal(counts, rcounts, err_msg=msg)
        np.testing.assert_array_equal(ids, rids, err_msg=msg)
        np.testing.assert_array_equal(dists.view(np.uint32), rdists.view(np.uint32), err_msg=msg)
        for key in ("expansions", "dist_level0", "dist_upper", "list_ints_read"):
            assert st[key] == rst[key], (msg, key, st[key], rst[key])
    idx.setBeamWaves(0)


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "dec100_cos", "gmm96_nd", "u10k_l2",
                                  "u67_l1", "u64_linf"])
@pytest.mark.parametrize("beam", [1, 2, 5, 16, 32, 64])
@pytest.mark.parametrize("gather", [0, 2])
def test_gpu_coop_beam_equals_model(name, beam, gather):
    # the cooperative kernel (one query per block of eight waves; bit marks set
    # by atomic ORs that also dedupe; per-wave sorted runs ranked by binary
    # search; the rebuild in rounds of 64 chunks) against the CPU model: ids,
    # distances and work counters identical, up to 2048 entries per step.
    # gather 0: two blocks per CU where LDS holds two queries (always, at these
    # sizes); gather 2: one block per CU with the wide distance path
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    oracle = O.OracleIndex(F.index_path(name))
    w = max(1, min(beam, 2048 // oracle.maxM0))
    q = F.results(name)["queries"]
    idx.setBeam(beam)
    idx.setBeamWaves(8)
    idx.setBeamGather(gather)
    idx.collectStats(True)
    for k, ef in [(10, 64), (100, 32), (10, 300), (1, 700)]:
        idx.setQueryTimeParams({"efSearch": ef, "algoType": "v1merge"})
        ids, dists, counts = idx.knnQueryBatchArrays(q, k)
        st = idx.lastStats()
        launch = idx.lastLaunch()
        two = gather == 0 and 163840 // launch["lds_bytes_per_query"] >= 2
        assert launch["kernel"] == ("hnsw_search_kernel_coop2" if two else "hnsw_search_kernel_coop")
        rids, rdists, rcounts, rst = oracle.search_beam(q, k, ef, w, num_threads=8, with_stats=True)
        msg = f"{name} beam={beam} {launch['kernel']} k={k} ef={ef}"
        print(msg, flush=True)
        np.testing.assert_array_equal(counts, rcounts, err_msg=msg)
        np.testing.assert_array_equal(ids, rids, err_msg=msg)
        np.testing.assert_array_equal(dists.view(np.uint32), rdists.view(np.uint32), err_msg=msg)
        for key in ("expansions", "dist_level0", "dist_upper", "list_ints_read"):
            assert st[key] == rst[key], (msg, key, st[key], rst[key])
    idx.setBeamGather(0)
    idx.setBeamWaves(0)
    idx.setBeam(0)


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "u10k_l2"])
@pytest.mark.parametrize("beam", [4, 16, 64])
@pytest.mark.parametrize("screen,gather", [(1, 0), (2, 0), (2, 2)])
def test_gpu_coop_screen_equals_model(name, beam, screen, gather):
    # the BF16 screen of the cooperative FP32 l2 search: rows, distances and
    # counters equal the FP32 beam model's; screen 2 checks every bound
    # against the FP32 distance on the device (a failure raises)
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    oracle = O.OracleIndex(F.index_path(name))
    q = F.results(name)["queries"]
    idx.setBeam(beam)
    idx.setBeamWaves(8)
    idx.setBeamGather(gather)
    idx.setScreen(screen)
    idx.collectStats(True)
    screened = 0
    for k, ef in [(10, 16), (10, 64), (1, 300)]:
        idx.setQueryTimeParams({"efSearch": ef, "algoType": "v1merge"})
        ids, dists, counts = idx.knnQueryBatchArrays(q, k)
        st = idx.lastStats()
        screened += st["screened"]
        assert st["screen_pass"] <= st["screened"] <= st["dist_level0"]
        rids, rdists, rcounts, rst = oracle.search_beam(q, k, ef, beam, num_threads=8,
                                                        with_stats=True)
        msg = f"{name} beam={beam} screen={screen} gather={gather} k={k} ef={ef}"
        print(msg, st["screened"], st["screen_pass"], flush=True)
        np.testing.assert_array_equal(counts, rcounts, err_msg=msg)
        np.testing.assert_array_equal(ids, rids, err_msg=msg)
        np.testing.assert_array_equal(dists.view(np.uint32), rdists.view(np.uint32), err_msg=msg)
        for key in ("expansions", "dist_level0", "dist_upper", "list_ints_read"):
            assert st[key] == rst[key], (msg, key, st[key], rst[key])
    assert screened > 0
    idx.setScreen(0)
    idx.setBeamGather(0)
    idx.setBeamWaves(0)
    idx.setBeam(0)


def test_gpu_bf16_on_integer_data_is_exact():
    # coordinates 0..255 are exact in BF16 and every partial sum is an exact
    # integer below 2^24, so BF16 traversal reproduces the FP32 results
    name = "hn128_l2"
    idx = load(name, F.index_path(name), "l2", 100)
    q = F.results(name)["queries"]
    ref = idx.knnQueryBatchArrays(q, 10)
    idx.setPrecision("bf16", 0)
    got = idx.knnQueryBatchArrays(q, 10)
    for x, y in zip(ref, got):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("storage", ["bf16", "fp8"])
@pytest.mark.parametrize("name", ["u128_l2", "dec100_cos", "gmm96_nd", "u67_l1", "u64_linf"])
def test_gpu_reduced_precision_with_rerank(name, storage):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"], 100)
    res = F.results(name)
    q = res["queries"]
    ex_ids, _, _ = idx.knnQueryBatchArrays(q, 10)
    idx.setPrecision(storage, 4)
    ids, dists, counts = idx.knnQueryBatchArrays(q, 10)
    oracle = O.OracleIndex(F.index_path(name))
    blk = oracle.level0_block().reshape(oracle.n, oracle.elem_bytes)
    vec = blk[:, 16:16 + 4 * oracle.dim].copy().view(np.float32)
    # reranked distances are the reference's FP32 distances, bit for bit
    for i in range(len(q)):
        qq = O.normalize(q[i]) if m["space"] == "cosinesimil" else q[i]
        for j in range(counts[i]):
            assert dists[i, j] == _exact_dist(oracle.func, qq, vec[ids[i, j]])
    x = F.base_vectors(name)
    gt, gd = O.brute_force(x, q, 100, m["space"])
    r_exact = F.recall_tie_aware(ex_ids, gt, gd, 10, rel=1e-5)
    r_low = F.recall_tie_aware(ids, gt, gd, 10, rel=1e-5)
    tol = 0.01 if storage == "bf16" else 0.05
    print(f"{name} {storage}: recall exact {r_exact:.4f} reduced+rerank {r_low:.4f}")
    assert r_low >= r_exact - tol

"""The device brute force against the reference's seq_search goldens
(method/seqsearch.cc:143-177, recorded by tests/golden/make_goldens.py) and
the FP32 distances of the oracle.  The matrix-core screen keeps kc candidates,
their distances are recomputed in the reference's FP32 summation order and
ranked by (distance, id); a query whose certificate fails is scanned again
exactly, so the answers are exact on any data."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import fixtures as F
from tests.gpu_util import load

pytestmark = pytest.mark.gpu


def _stored_vectors(name):
    oracle = O.OracleIndex(F.index_path(name))
    blk = oracle.level0_block().reshape(oracle.n, oracle.elem_bytes)
    return oracle, blk[:, 16:16 + 4 * oracle.dim].copy().view(np.float32)


@pytest.mark.parametrize("name", ["u128_l2", "u100_l2", "hn128_l2", "u10k_l2", "tiny_l2",
                                  "gmm96_nd"])
def test_seq_search_equals_reference_goldens(name):
    # l2 (sqrt of the SSE-order sum) and negdotprod: ids and FP32 distance
    # bits equal the reference's seq_search; exactly tied distances are
    # compared as sets (the reference orders them by Object address)
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    res = F.results(name)
    q = res["queries"]
    k = min(100, m["n"])
    ids, dists, counts = idx.seqSearch(q, k)
    assert np.all(counts == k)
    gi, gd = F.canon(res["bf_ids"][:, :k], res["bf_dists"][:, :k])
    ci, cd = F.canon(ids, dists)
    np.testing.assert_array_equal(cd.view(np.uint32), gd.view(np.uint32))
    np.testing.assert_array_equal(ci, gi)
    # rows come back in (distance, id) order
    np.testing.assert_array_equal(ids, ci)
    print(f"{name}: {idx.lastBruteForceRescans()} of {len(q)} queries took the exact scan")


def test_seq_search_cosine_over_normalised_vectors():
    # the optimized cosine index stores NormalizeVect'ed vectors, the
    # reference's seq_search ran over the raw ones: the same neighbours up to
    # the rounding of the normalisation
    name = "dec100_cos"
    idx = load(name, F.index_path(name), "cosinesimil")
    res = F.results(name)
    ids, dists, counts = idx.seqSearch(res["queries"], 100)
    assert np.max(np.abs(dists - res["bf_dists"])) < 2e-6
    r = F.recall_tie_aware(ids, res["bf_ids"], res["bf_dists"], 100, rel=2e-6)
    assert r > 0.999, r


@pytest.mark.parametrize("name", ["u128_l2", "hn128_l2", "dec100_cos", "gmm96_nd"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_bruteforce_is_exact_in_the_search_distance(name, k):
    # every returned distance is the search kernel's FP32 value (the oracle's
    # AVX-order function), and nothing outside the answer ranks before its
    # last entry: checked against all points for a sample of queries
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    q = F.results(name)["queries"]
    ids, dists, counts = idx.bruteForce(q, k)
    assert np.all(counts == k)
    oracle, vec = _stored_vectors(name)
    ext = np.arange(oracle.n)  # fixtures store object id = row
    for i in range(0, len(q), 20):
        qq = O.normalize(q[i]) if m["space"] == "cosinesimil" else q[i]
        full = np.array([O.dist(oracle.func, qq, vec[j]) for j in range(oracle.n)], np.float32)
        order = np.lexsort((ext, full))[:k]
        np.testing.assert_array_equal(ids[i], order, err_msg=f"{name} query {i}")
        np.testing.assert_array_equal(dists[i].view(np.uint32), full[order].view(np.uint32))


def test_bruteforce_certificate_falls_back_to_the_exact_scan():
    # 10 distinct points, each stored 100 times: the k-th and the kc-th screen
    # distances tie, no screen can decide, and the exact scan answers
    import nmslib_amd as nmslib
    from tests.gpu_util import require_gpu
    require_gpu()
    rng = np.random.default_rng(5)
    base = rng.random((10, 32), dtype=np.float32)
    x = np.repeat(base, 100, axis=0)
    idx = nmslib.init(method="hnsw", space="l2")
    idx.addDataPointBatch(x)
    idx.createIndex({"M": 8, "efConstruction": 50, "indexThreadQty": 1})
    q = rng.random((16, 32), dtype=np.float32)
    ids, dists, counts = idx.bruteForce(q, 10)
    assert idx.lastBruteForceRescans() == len(q)
    func = idx.info()["dist_func_type"]
    for i in range(len(q)):
        full = np.array([O.dist(func, q[i], v) for v in x], np.float32)
        order = np.lexsort((np.arange(len(x)), full))[:10]
        np.testing.assert_array_equal(ids[i], order)
        np.testing.assert_array_equal(dists[i].view(np.uint32), full[order].view(np.uint32))


def test_bruteforce_integer_data_needs_no_rescan():
    # coordinates 0..255: the screen is exact, every certificate holds
    name = "hn128_l2"
    idx = load(name, F.index_path(name), "l2")
    q = F.results(name)["queries"]
    ids, dists, counts = idx.bruteForce(q, 100)
    assert idx.lastBruteForceRescans() == 0
    gt, gd = O.brute_force(F.base_vectors(name), q, 100, "l2")
    np.testing.assert_array_equal(ids, gt)
    np.testing.assert_array_equal(dists, gd.astype(np.float32))

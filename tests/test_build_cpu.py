"""Index construction (createIndex) against the reference on the CPU.

The builder runs on the host, so these tests need no GPU.  With one build
thread the graph must equal the reference's exactly:

* the fixtures bld_<name>.npz (tests/golden/make_build_goldens.py) were each
  built by the reference in a fresh process; each is rebuilt here on a fresh
  thread, whose level generator likewise starts from seed 0, and the saved
  file must be byte-identical;
* the query fixtures idx_<name>.npz (tests/golden/make_goldens.py) were built
  one after another in one process, the first seven in one run and the l1 /
  linf pair in a second, so the level generator carried on from build to
  build; replaying each run's builds in order on one fresh thread must give
  the same graphs (compared list by list: the reference leaves the unused
  slots of upper-level lists uninitialised).

Several build threads make the graph depend on scheduling, as in the
reference, so that mode is checked by invariants and by recall through the
oracle's search of the saved file.
"""
import json
import os
import tempfile
import threading

import numpy as np
import pytest

import nmslib_amd as nm
from nmslib_amd import datasets as D
from nmslib_amd import indexpack
from oracle import oracle as O
from tests import fixtures as F

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BUILD_CASES = json.load(open(os.path.join(GOLDEN, "build_cases.json")))
_TMP = tempfile.mkdtemp(prefix="nmsamd_build_")


def build_on_fresh_thread(jobs):
    """Runs addDataPointBatch + createIndex(params + indexThreadQty=1) +
    saveIndex for each (space, x, params, out) in order, on one new thread."""
    err = []

    def run():
        try:
            for space, x, params, out in jobs:
                idx = nm.init(method="hnsw", space=space)
                idx.addDataPointBatch(x)
                idx.createIndex(params.split(",") + ["indexThreadQty=1"])
                idx.saveIndex(out)
        except Exception as e:  # re-raised on the test's thread
            err.append(e)

    t = threading.Thread(target=run)
    t.start()
    t.join()
    if err:
        raise err[0]


def assert_same_graph(a, b):
    """Header, level-0 block and upper-level lists (count and ids) equal."""
    A, B = indexpack.read_index(a), indexpack.read_index(b)
    assert A["header"] == B["header"]
    bad = np.nonzero(np.any(A["block"] != B["block"], axis=1))[0]
    assert bad.size == 0, f"{bad.size} level-0 elements differ, first {bad[:5]}"
    np.testing.assert_array_equal(A["levels"], B["levels"])
    per = A["header"]["maxM"] + 1
    ua, ub = A["upper"].reshape(-1, per), B["upper"].reshape(-1, per)
    np.testing.assert_array_equal(ua[:, 0], ub[:, 0])
    for r in range(ua.shape[0]):
        c = ua[r, 0]
        np.testing.assert_array_equal(ua[r, 1:1 + c], ub[r, 1:1 + c])


@pytest.mark.parametrize("name", sorted(BUILD_CASES))
def test_build_is_byte_identical_to_reference(name):
    m = BUILD_CASES[name]
    rep = m.get("repeat", 1)
    x = np.repeat(getattr(D, m["generator"])(m["n"] // rep, m["dim"], m["seed"]), rep, axis=0)
    ours = os.path.join(_TMP, f"{name}.ours")
    build_on_fresh_thread([(m["space"], x, m["params"], ours)])
    ref = os.path.join(_TMP, f"{name}.ref")
    indexpack.unpack(indexpack.load_pack(os.path.join(GOLDEN, f"bld_{name}.npz")), x, ref)
    assert open(ours, "rb").read() == open(ref, "rb").read()


# the two runs of make_goldens.py that wrote the query fixtures, in build order
GOLDEN_RUNS = [["u128_l2", "u10k_l2", "hn128_l2", "u100_l2", "dec100_cos", "gmm96_nd", "tiny_l2"],
               ["u67_l1", "u64_linf"]]


@pytest.mark.parametrize("run", range(len(GOLDEN_RUNS)))
def test_build_replays_the_query_fixtures(run):
    names = GOLDEN_RUNS[run]
    jobs = []
    for name in names:
        m = F.META[name]
        jobs.append((m["space"], F.base_vectors(name), m["params"],
                     os.path.join(_TMP, f"{name}.replay")))
    build_on_fresh_thread(jobs)
    for name, job in zip(names, jobs):
        assert_same_graph(job[3], F.index_path(name))


def test_add_data_points_ids_and_len(tmp_path):
    x = D.uniform(300, 8, 1)
    idx = nm.init(method="hnsw", space="l2")
    assert idx.addDataPointBatch(x[:100]) == 100
    assert idx.addDataPointBatch(x[100:250], ids=np.arange(1000, 1150)) == 150
    assert idx.addDataPointBatch([list(v) for v in x[250:299]]) == 49
    assert idx.addDataPoint(77, x[299]) == 299
    assert len(idx) == 300
    idx.createIndex({"M": 6, "efConstruction": 30, "indexThreadQty": 2})
    path = str(tmp_path / "ids.hnsw")
    idx.saveIndex(path)
    ix = indexpack.read_index(path)
    head = ix["block"][:, :16].copy().view(np.int32)
    # readObjectVector numbers rows within each batch when no ids are given
    want = np.concatenate([np.arange(100), np.arange(1000, 1150), np.arange(49), [77]])
    np.testing.assert_array_equal(head[:, 0], want)
    assert np.all(head[:, 1] == -1) and np.all(head[:, 2] == 4 * 8) and np.all(head[:, 3] == 0)
    vec = ix["block"][:, 16:16 + 32].copy().view(np.float32)
    np.testing.assert_array_equal(vec, x)
    assert ix["header"]["func"] == 2  # L2SqrExt: 8 is not a multiple of 16


def _graph_invariants(path):
    ix = indexpack.read_index(path)
    h = ix["header"]
    n, maxM, maxM0 = h["n"], h["maxM"], h["maxM0"]
    l0 = ix["block"][:, h["off_l0"]:h["off_l0"] + 4 * (maxM0 + 1)].copy().view(np.int32)
    for i in range(n):
        c = l0[i, 0]
        assert 1 <= c <= maxM0
        lst = l0[i, 1:1 + c]
        assert np.all((lst >= 0) & (lst < n)) and i not in lst
        assert len(set(lst.tolist())) == c
    levels = ix["levels"]
    assert levels.max() == h["maxlevel"] and levels[h["enter"]] == h["maxlevel"]
    up = ix["upper"].reshape(-1, maxM + 1)
    assert np.all((up[:, 0] >= 0) & (up[:, 0] <= maxM))
    row = 0
    for i in np.nonzero(levels)[0]:
        for lev in range(1, levels[i] + 1):
            lst = up[row, 1:1 + up[row, 0]]
            assert np.all(levels[lst] >= lev) and i not in lst
            row += 1


def test_multithreaded_build_invariants_and_recall():
    x = D.uniform(4000, 32, 31)
    q = D.uniform(200, 32, 32)
    gt, gd = O.brute_force(x, q, 10, "l2")
    rec = {}
    for threads in (1, 8):
        path = os.path.join(_TMP, f"mt{threads}.hnsw")
        idx = nm.init(method="hnsw", space="l2")
        idx.addDataPointBatch(x)
        idx.createIndex({"M": 12, "efConstruction": 100, "indexThreadQty": threads})
        idx.saveIndex(path)
        _graph_invariants(path)
        ids, _, _ = O.OracleIndex(path).search(q, 10, 64, "v1merge")
        rec[threads] = F.recall_tie_aware(ids, gt, gd, 10)
    assert rec[1] >= 0.95 and rec[8] >= 0.95
    assert abs(rec[8] - rec[1]) <= 0.02, rec


@pytest.mark.parametrize("params, exc, msg", [
    ({"M": 8, "foo": 1}, RuntimeError, "Unknown parameters found!"),
    (["M=abc"], RuntimeError, "Failed to convert value 'abc' from type: m"),
    (["M"], RuntimeError, "Wrong format of an argument"),
    ({"delaunay_type": 3, "indexThreadQty": 2}, ValueError, "delaunay_type 3 needs indexThreadQty=1"),
    ({"delaunay_type": 4}, ValueError, "delaunay_type should be"),
    ({"post": 3}, ValueError, "post"),
    ({"M": 40}, ValueError, "device limit"),
    ({"M": 1}, ValueError, "mult"),
])
def test_create_index_rejects_bad_parameters(params, exc, msg):
    idx = nm.init(method="hnsw", space="l2")
    idx.addDataPointBatch(D.uniform(50, 4, 2))
    with pytest.raises(exc, match=msg):
        idx.createIndex(params)


def test_create_index_input_errors():
    idx = nm.init(method="hnsw", space="cosinesimil")
    with pytest.raises(ValueError, match="no data points"):
        idx.createIndex()
    idx.addDataPointBatch(D.uniform(20, 4, 2))
    idx.addDataPointBatch(D.uniform(20, 5, 3))
    with pytest.raises(RuntimeError, match="differ in dimension"):
        idx.createIndex()
    with pytest.raises(RuntimeError, match="2d array"):
        idx.addDataPointBatch(np.zeros(4, np.float32))
    with pytest.raises(ValueError, match="one id per data point"):
        idx.addDataPointBatch(np.zeros((3, 4), np.float32), ids=[1, 2])

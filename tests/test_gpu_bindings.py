"""The reference's own dense-index binding tests
(python_bindings/tests/bindings_test.py:117-193, DenseIndexTestMixin: knnQuery
with row- and column-major input, knnQueryBatch with custom ids, save and
reload with and without data) run on the drop-in: the index is built by
createIndex on the host and searched on the device."""
import os
import shutil
import tempfile

import numpy as np
import numpy.testing as npt
import pytest

import nmslib_amd as nmslib
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def get_exact_cosine(row, data, N=10):
    scores = data.dot(row) / np.linalg.norm(data, axis=-1)
    best = np.argpartition(scores, -N)[-N:]
    return sorted(zip(best, scores[best] / np.linalg.norm(row)), key=lambda x: -x[1])


def get_hitrate(ground_truth, ids):
    return len(set(i for i, _ in ground_truth).intersection(ids))


def _get_index(space="cosinesimil"):
    return nmslib.init(method="hnsw", space=space)


def assert_allclose(orig, comp):  # ties resorted into the first result's order
    qty = len(orig[0])
    ids2dist = {comp[0][k]: comp[1][k] for k in range(qty)}
    assert len(comp[0]) == qty and all(i in ids2dist for i in orig[0])
    npt.assert_allclose(orig, (list(orig[0]), [ids2dist[i] for i in orig[0]]), rtol=1e-7, atol=1e-7)


def test_knn_query():
    require_gpu()
    np.random.seed(23)
    data = np.asfortranarray(np.random.randn(1000, 10).astype(np.float32))
    index = _get_index()
    index.addDataPointBatch(data)
    index.createIndex()
    query = data[0]
    ids, distances = index.knnQuery(query, k=10)
    assert get_hitrate(get_exact_cosine(query, data), ids) >= 5
    query = np.array([1, 1, 1, 1, 1, 1, 1, 1, 1, 1.])
    ids, distances = index.knnQuery(query, k=10)
    assert get_hitrate(get_exact_cosine(query, data), ids) >= 5


def test_knn_query_batch():
    require_gpu()
    np.random.seed(23)
    data = np.random.randn(1000, 10).astype(np.float32)
    index = _get_index()
    index.addDataPointBatch(data)
    index.createIndex()
    queries = data[:10]
    for q in (queries, np.asfortranarray(queries)):
        for query, (ids, distances) in zip(q, index.knnQueryBatch(q, k=10)):
            assert get_hitrate(get_exact_cosine(query, data), ids) >= 5
    index = _get_index()
    index.addDataPointBatch(data, ids=np.arange(data.shape[0]) ** 2)
    index.createIndex()
    for query, (ids, distances) in zip(queries, index.knnQueryBatch(queries, k=10)):
        ids = np.sqrt(ids).astype(int)
        assert get_hitrate(get_exact_cosine(query, data), ids) >= 5


@pytest.mark.parametrize("space, params", [("cosinesimil", None), ("l2", None),
                                           ("l2", {"skip_optimized_index": 1})])
def test_reload_index(space, params):
    require_gpu()
    np.random.seed(23)
    data = np.random.randn(1000, 10).astype(np.float32)
    original = _get_index(space)
    original.addDataPointBatch(data)
    original.createIndex(params)
    temp_dir = tempfile.mkdtemp()
    pref = os.path.join(temp_dir, "index")
    for save_data in [0, 1]:
        original.saveIndex(pref, save_data=save_data)
        reloaded = _get_index(space)
        if save_data == 0:
            reloaded.addDataPointBatch(data)
        reloaded.loadIndex(pref, load_data=save_data)
        assert_allclose(original.knnQuery(data[0]), reloaded.knnQuery(data[0]))
    shutil.rmtree(temp_dir)

"""The Thrift QueryService front-end on the device: knnQuery and
knnQueryBatch answered over a socket equal knnQueryBatch's rows (ids; FP32
distances widened to the IDL's double), from several concurrent client
connections whose queries the dispatcher batches together, with externId and
the object text as QueryService_server.cpp:320-466 builds them."""
import threading

import numpy as np
import pytest

import nmslib_amd as nmslib
from nmslib_amd import query_server as qs
from tests import fixtures as F
from tests.gpu_util import require_gpu
from tests.test_query_server_cpu import _Client

pytestmark = pytest.mark.gpu


def _entries(res):
    """[(id, dist, externId, obj)] of a knnQuery reply's success field."""
    assert 0 in res, res
    out = []
    for e in res[0][1]:
        out.append((e[1][1], e[2][1], e.get(3, (0, None))[1], e.get(4, (0, None))[1]))
    return out


@pytest.fixture(scope="module")
def served():
    require_gpu()
    name = "u128_l2"
    x = F.base_vectors(name)
    idx = nmslib.init(method="hnsw", space="l2")
    idx.addDataPointBatch(x)          # the data points (for retObj)
    idx.loadIndex(F.index_path(name))  # load_data=False keeps them (nmslib.cc:103-107)
    extern = [f"doc-{i}" for i in range(len(x))]
    srv = qs.serve(idx, 0, extern_ids=extern, lanes=2, max_queries=256, max_k=100)
    c = _Client(srv.server_address[1])
    mtype, res = c.call("setQueryTimeParams", [(1, qs.T_STRING, b"efSearch=100,algoType=v1merge")])
    assert res == {}
    c.close()
    yield srv, idx, x, extern
    srv.shutdown()
    srv.server_close()
    srv.handler.close()


def test_knn_query_over_the_wire_equals_the_batch_rows(served):
    srv, idx, x, extern = served
    q = F.results("u128_l2")["queries"]
    ref_ids, ref_d, ref_c = idx.knnQueryBatchArrays(q, 10)
    errors = []

    def client(rows, k):
        try:
            c = _Client(srv.server_address[1])
            for i in rows:
                mtype, res = c.call("knnQuery", [(1, qs.T_I32, k),
                                                 (2, qs.T_STRING, qs.vector_text(q[i]).encode()),
                                                 (3, qs.T_BOOL, True), (4, qs.T_BOOL, i % 5 == 0)])
                got = _entries(res)
                assert [g[0] for g in got] == ref_ids[i, :k].tolist()
                assert [g[1] for g in got] == [float(v) for v in ref_d[i, :k]]
                assert [g[2] for g in got] == [extern[j].encode() for j in ref_ids[i, :k]]
                if i % 5 == 0:
                    assert got[0][3] == qs.vector_text(x[ref_ids[i, 0]]).encode()
            c.close()
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    # query text round trip: %.9g reproduces every float32 exactly
    assert np.array_equal(qs.parse_vector(qs.vector_text(q[0]).encode()), q[0])
    threads = [threading.Thread(target=client, args=(range(t, 120, 4), 10)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[0]


def test_knn_query_batch_over_the_wire(served):
    srv, idx, x, extern = served
    q = F.results("u128_l2")["queries"][:50]
    ref_ids, ref_d, ref_c = idx.knnQueryBatchArrays(q, 7)
    c = _Client(srv.server_address[1])
    mtype, res = c.call("knnQueryBatch", [(1, qs.T_I32, 7),
                                          (2, qs.T_LIST, [qs.vector_text(v).encode() for v in q]),
                                          (3, qs.T_BOOL, False), (4, qs.T_BOOL, False),
                                          (5, qs.T_I32, 0)])
    rows = res[0][1]
    assert len(rows) == len(q)
    for i, row in enumerate(rows):
        assert [e[1][1] for e in row] == ref_ids[i].tolist()
        assert [e[2][1] for e in row] == [float(v) for v in ref_d[i]]
        assert all(3 not in e for e in row)  # no externId unless asked
    c.close()

"""The no-allocation device entry point (nmsamd_knn_query_batch_device_noalloc,
no reference counterpart) against nmsamd_knn_query_batch_device through the C
ABI: the same ids, distances and counts, on two calls in a row whose second
is smaller, for exact V1Merge, SearchOld, beam mode and the cooperative beam
kernel (bit marks: slots of n / 8 bytes)."""
import numpy as np
import pytest

from tests import fixtures as F
from tests.gpu_util import load

pytestmark = pytest.mark.gpu


def _run(idx, entry, q, k):
    import torch
    dev = torch.device("cuda:0")
    dq = torch.from_numpy(np.ascontiguousarray(q, np.float32)).to(dev)
    ids = torch.full((len(q), k), -7, dtype=torch.int32, device=dev)
    dists = torch.full((len(q), k), -7.0, dtype=torch.float32, device=dev)
    counts = torch.full((len(q),), -7, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    entry(dq.data_ptr(), len(q), k, ids.data_ptr(), dists.data_ptr(), counts.data_ptr(), stream)
    torch.cuda.synchronize()
    return ids.cpu().numpy(), dists.cpu().numpy(), counts.cpu().numpy()


@pytest.mark.parametrize("name,params,beam,waves,k", [
    ("u128_l2", {"efSearch": 100, "algoType": "v1merge"}, 0, 0, 10),
    ("u10k_l2", {"efSearch": 64, "algoType": "old"}, 0, 0, 10),
    ("dec100_cos", {"efSearch": 300, "algoType": "v1merge"}, 8, 0, 100),
    ("u10k_l2", {"efSearch": 300, "algoType": "v1merge"}, 64, 8, 10),
])
def test_no_alloc_entry_equals_device_entry(name, params, beam, waves, k):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    idx.setQueryTimeParams(params)
    idx.setBeam(beam)
    idx.setBeamWaves(waves)
    q = F.results(name)["queries"]
    idx.knnQueryBatchArrays(q[:4], k)  # the index reaches the device
    for rows in (slice(0, len(q)), slice(3, 3 + len(q) // 3)):  # the second call is smaller
        qq = q[rows]
        got = _run(idx, idx.knnQueryBatchDeviceNoAlloc, qq, k)
        ref = _run(idx, idx.knnQueryBatchDevice, qq, k)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        assert np.all(got[2] == min(k, m["n"]))
    idx.setBeamWaves(0)
    idx.setBeam(0)

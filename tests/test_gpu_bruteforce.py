"""The device brute force (MFMA) against the FP64 oracle and the reference's
seq_search goldens: exact on integer-valued data (ids including tie order by
id, distances bit-exact), within float rounding elsewhere."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import fixtures as F
from tests.gpu_util import load

pytestmark = pytest.mark.gpu


def test_bruteforce_integer_data_is_exact():
    # coordinates 0..255: every dot product and norm is an exact integer < 2^24
    name = "hn128_l2"
    idx = load(name, F.index_path(name), "l2")
    q = F.results(name)["queries"]
    ids, dists, counts = idx.bruteForce(q, 100)
    gt, gd = O.brute_force(F.base_vectors(name), q, 100, "l2")
    assert np.all(counts == 100)
    np.testing.assert_array_equal(ids, gt)
    np.testing.assert_array_equal(dists, gd.astype(np.float32))


@pytest.mark.parametrize("name", ["u128_l2", "u100_l2", "dec100_cos", "gmm96_nd", "tiny_l2"])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_bruteforce_matches_fp64_within_rounding(name, k):
    m = F.META[name]
    idx = load(name, F.index_path(name), m["space"])
    q = F.results(name)["queries"]
    k = min(k, m["n"])
    ids, dists, counts = idx.bruteForce(q, k)
    assert np.all(counts == k)
    gt, gd = O.brute_force(F.base_vectors(name), q, k, m["space"])
    # distances agree with FP64 up to the rounding of the |x|^2+|q|^2-2q.x form
    scale = np.maximum(1.0, np.abs(gd))
    assert np.max(np.abs(dists - gd) / scale) < 2e-5
    # ids agree except where two exact distances are within that rounding
    x64 = F.base_vectors(name).astype(np.float64)
    for i in np.nonzero(np.any(ids != gt, axis=1))[0]:
        for j in np.nonzero(ids[i] != gt[i])[0]:
            assert abs(dists[i, j] - gd[i, j]) <= 2e-5 * max(1.0, abs(gd[i, j]))
    assert np.mean(ids == gt) > 0.99
    # rows sorted, ids unique
    assert np.all(np.diff(dists, axis=1) >= 0)
    for row in ids:
        assert len(set(row.tolist())) == k

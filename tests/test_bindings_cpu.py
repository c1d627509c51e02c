"""The data-point surface of the binding on the host: __getitem__,
getDistance (the space's own distance, python_bindings/nmslib.cc:348-353) and
the legacy module-level API (nmslib.cc:694-789), after the reference's
python_bindings/tests/legacy_test.py DenseTests (there on method
small_world_rand; the data handling does not depend on the method)."""
import numpy as np
import numpy.testing as nt
import pytest

import nmslib_amd as nmslib


def _index(space="cosinesimil"):
    return nmslib.init(space, [], "hnsw", nmslib.DataType.DENSE_VECTOR, nmslib.DistType.FLOAT)


def test_legacy_add_points():
    index = _index()
    assert nmslib.addDataPoint(index, 1000, [0.5, 0.3, 0.4]) == 0
    assert nmslib.addDataPoint(index, 1001, [0.5, 0.3, 0.4]) == 1
    assert nmslib.getDataPointQty(index) == 2
    assert nmslib.getDataPoint(index, 1) == pytest.approx([0.5, 0.3, 0.4])


@pytest.mark.parametrize("ids, data", [
    ([0, 1, 2], [[0.34, 0.54], [0.55, 0.52], [0.21, 0.68]]),
    (np.array([0, 1, 2], dtype=np.int32), [[0.34, 0.54], [0.55, 0.52], [0.21, 0.68]]),
    (np.array([0, 1, 2], dtype=np.int32), np.array([[0.34, 0.54], [0.55, 0.52], [0.21, 0.68]])),
])
def test_legacy_add_points_batch_rejects_non_float32_arrays(ids, data):
    with pytest.raises(ValueError):
        nmslib.addDataPointBatch(_index(), ids, data)


def test_legacy_add_points_batch_returns_positions():
    index = _index()
    data = np.array([[0.34, 0.54], [0.55, 0.52], [0.21, 0.68]], dtype=np.float32)
    positions = nmslib.addDataPointBatch(index, np.array([0, 1, 2], dtype=np.int32), data)
    nt.assert_array_equal(np.array([0, 1, 2], dtype=np.int32), positions)
    positions = nmslib.addDataPointBatch(index, np.array([7, 8, 9], dtype=np.int32), data)
    nt.assert_array_equal(np.array([3, 4, 5], dtype=np.int32), positions)


def test_getitem_and_get_distance():
    x = np.array([[3.0, 4.0, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0, 0.0, 0.0],
                  [1.0, 2.0, 2.0, 0.0, 0.0], [-3.0, -4.0, 0.0, 0.0, -1.0]], np.float32)
    cos = _index("cosinesimil")
    cos.addDataPointBatch(x)
    assert cos[2] == [1.0, 2.0, 2.0, 0.0, 0.0]
    with pytest.raises(IndexError):
        cos[4]
    # a zero vector is at distance 1 from everything, itself included
    # (distcomp_scalar.cc:153-159, python_bindings/tests/bindings_test.py:73-91)
    assert cos.getDistance(1, 1) == 1.0 and cos.getDistance(0, 1) == 1.0
    assert cos.getDistance(0, 3) == 2.0 and cos.getDistance(0, 0) == pytest.approx(0.0, abs=1e-7)
    l2 = _index("l2")
    l2.addDataPointBatch(x)
    assert l2.getDistance(0, 2) == np.float32(np.sqrt(np.float32(4 + 4 + 4 + 0 + 1)))  # sqrt, not squared
    nd = _index("negdotprod")
    nd.addDataPointBatch(x)
    assert nd.getDistance(0, 2) == -11.0
    l1 = _index("l1")
    l1.addDataPointBatch(x)
    assert l1.getDistance(0, 3) == 16.0
    linf = _index("linf")
    linf.addDataPointBatch(x)
    assert linf.getDistance(0, 3) == 8.0


def test_import_nmslib_alias_is_the_device_module():
    # `import nmslib` code is a drop-in: the reference module's names resolve
    # to this library's (python_bindings/nmslib.cc:458-789)
    import nmslib as alias
    assert alias.init is nmslib.init
    assert alias.DataType.DENSE_VECTOR == nmslib.DataType.DENSE_VECTOR
    idx = alias.init(method="hnsw", space="l2")
    assert isinstance(idx, nmslib.FloatIndex)
    assert alias.addDataPoint(idx, 7, [0.1, 0.2]) == 0
    assert alias.getDataPointQty(idx) == 1

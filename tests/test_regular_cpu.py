"""The regular (non-optimized) index format on the host: load it over the
data points, save it back byte-identical, build it with
skip_optimized_index=1 byte-identical to the reference, and the errors the
reference raises (hnsw.cc:810-841, 943-991).  Each fixture was built by the
reference single-threaded in a fresh process, so it is rebuilt here on a
fresh thread (whose level generator also starts from seed 0)."""
import os
import tempfile
import threading

import numpy as np
import pytest

import nmslib_amd as nm
from tests import regular_fixtures as R

_TMP = tempfile.mkdtemp(prefix="nmsamd_reg_")


@pytest.mark.parametrize("name", R.NAMES)
def test_regular_index_loads_over_added_data_and_saves_back(name):
    m = R.META[name]
    path = os.path.join(_TMP, f"{name}.reg")
    R.write_index(name, path)
    idx = nm.init(method="hnsw", space=m["space"])
    idx.addDataPointBatch(R.base_vectors(name))
    idx.loadIndex(path)  # load_data=False keeps the points added before
    info = idx.info()
    assert info["n"] == m["n"] and info["dim"] == m["dim"]
    out = os.path.join(_TMP, f"{name}.back")
    idx.saveIndex(out, save_data=True)
    assert open(out, "rb").read() == open(path, "rb").read()
    # the data file written with it loads the index into a fresh handle
    again = nm.init(method="hnsw", space=m["space"])
    again.loadIndex(out, load_data=True)
    assert len(again) == m["n"]
    out2 = os.path.join(_TMP, f"{name}.back2")
    again.saveIndex(out2)
    assert open(out2, "rb").read() == open(path, "rb").read()


@pytest.mark.parametrize("name", R.NAMES)
def test_regular_build_is_byte_identical_to_reference(name):
    m = R.META[name]
    x = R.base_vectors(name)
    out = os.path.join(_TMP, f"{name}.built")
    err = []

    def run():
        try:
            idx = nm.init(method="hnsw", space=m["space"])
            idx.addDataPointBatch(x)
            idx.createIndex(m["params"].split(",") + ["skip_optimized_index=1", "indexThreadQty=1"])
            idx.saveIndex(out)
        except Exception as e:
            err.append(e)

    t = threading.Thread(target=run)
    t.start()
    t.join()
    if err:
        raise err[0]
    assert np.array_equal(np.fromfile(out, np.uint8), R.fixture(name)["index"])


def test_regular_index_needs_its_data():
    name = "r_u30_l2"
    path = os.path.join(_TMP, "nodata.reg")
    R.write_index(name, path)
    idx = nm.init(method="hnsw", space="l2")
    with pytest.raises(RuntimeError, match="doesn't match the number of data points 0! "
                                           "Did you forget to re-load data?"):
        idx.loadIndex(path)
    wrong = nm.init(method="hnsw", space="cosinesimil")
    wrong.addDataPointBatch(R.base_vectors(name))
    wrong.loadIndex(path)  # the file does not name its space: it loads over any data
    assert wrong.info()["dist_func_type"] == 8

"""``import nmslib`` for the dense-float HNSW subset this library serves.

The names of the reference's pybind11 module (python_bindings/nmslib.cc:
458-789: ``init``, ``DataType``, ``DistType``, the index object and the legacy
module-level functions) resolve to ``nmslib_amd``'s, so code written against
``nmslib`` runs unchanged on the device path when this directory is on
``sys.path`` (and the reference module is not installed ahead of it)."""
import nmslib_amd as _impl
from nmslib_amd import (DataType, DistType, FloatIndex, addDataPoint,  # noqa: F401
                        addDataPointBatch, createIndex, freeIndex, getDataPoint,
                        getDataPointQty, getDistance, init, knnQuery, knnQueryBatch,
                        loadIndex, saveIndex, setQueryTimeParams)

__all__ = ["init", "DataType", "DistType", "FloatIndex", "addDataPoint", "addDataPointBatch",
           "createIndex", "freeIndex", "getDataPoint", "getDataPointQty", "getDistance",
           "knnQuery", "knnQueryBatch", "loadIndex", "saveIndex", "setQueryTimeParams"]
__version__ = getattr(_impl, "__version__", "0")

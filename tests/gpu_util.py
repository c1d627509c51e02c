"""Helpers of the GPU tests: a device is required and the HIP library must be
the one answering (the product has no CPU fallback)."""
import functools

import pytest

import nmslib_amd as nmslib


@functools.lru_cache(maxsize=None)
def have_gpu():
    try:
        import torch  # noqa: F401  (only to count devices without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def require_gpu():
    if not have_gpu():
        pytest.skip("no HIP device visible")


def load(name, path, space, ef=None):
    require_gpu()
    idx = nmslib.init(method="hnsw", space=space)
    idx.loadIndex(path)
    if ef is not None:
        idx.setQueryTimeParams({"efSearch": ef})
    return idx

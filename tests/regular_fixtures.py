"""Loads the regular-format fixtures reg_<name>.npz of tests/golden/ (written
from the reference by tests/golden/make_regular_goldens.py)."""
from __future__ import annotations

import functools
import json
import os

import numpy as np

from nmslib_amd import datasets as D

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "regular_cases.json")))
NAMES = sorted(META)


def base_vectors(name):
    m = META[name]
    return getattr(D, m["generator"])(m["n"], m["dim"], m["seed"])


@functools.lru_cache(maxsize=None)
def fixture(name):
    with np.load(os.path.join(GOLDEN, f"reg_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def write_index(name, path):
    """The reference's regular index file (no vectors in it)."""
    fixture(name)["index"].tofile(path)


def runs(name):
    return [tuple(r) for r in META[name]["runs"]]

"""CPU-side checks of the product boundary: the HIP library loads, exports
every entry point include/nmslib_amd.h declares, and the host logic that runs
without a device (init, parameter parsing, error mapping) behaves as the
reference binding does."""
import os
import re

import pytest

import nmslib_amd as nmslib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "nmslib_amd.h")).read()
    return sorted(set(re.findall(r"\b(nmsamd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = nmslib.load_library()
    syms = header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    from nmslib_amd import serving
    assert set(syms) == set(nmslib.SIGNATURES) | set(serving.SERVER_SIGNATURES)


def test_init_spaces_and_methods():
    for sp in ("l2", "cosinesimil", "negdotprod"):
        idx = nmslib.init(method="hnsw", space=sp)
        assert "hnsw" in repr(idx)
    with pytest.raises(RuntimeError, match="space foo is not defined"):
        nmslib.init(space="foo")
    with pytest.raises(RuntimeError, match="method sw-graph"):
        nmslib.init(space="l2", method="sw-graph")
    with pytest.raises(ValueError):
        nmslib.init(space="l2", data_type=nmslib.DataType.SPARSE_VECTOR)


def test_query_time_params_semantics():
    # Hnsw::SetQueryTimeParams (hnsw.cc:472-507)
    idx = nmslib.init(space="l2")
    idx.setQueryTimeParams({"efSearch": 77})
    assert idx.info()["ef"] == 77
    idx.setQueryTimeParams(["ef=33", "algoType=v1merge"])
    assert idx.info()["ef"] == 33
    idx.setQueryTimeParams()  # None resets to the default ef = 20
    assert idx.info()["ef"] == 20
    with pytest.raises(RuntimeError, match="synonyms"):
        idx.setQueryTimeParams({"ef": 10, "efSearch": 10})
    with pytest.raises(RuntimeError, match="Unknown parameters found!"):
        idx.setQueryTimeParams({"efConstruction": 10})
    with pytest.raises(RuntimeError, match="algoType should be one of"):
        idx.setQueryTimeParams({"algoType": "fast"})
    with pytest.raises(RuntimeError, match="Failed to convert value 'x'"):
        idx.setQueryTimeParams({"ef": "x"})
    with pytest.raises(RuntimeError, match="Wrong format"):
        idx.setQueryTimeParams(["ef"])
    with pytest.raises(ValueError, match="Unknown type for parameter"):
        idx.setQueryTimeParams({"ef": [1]})
    idx.setQueryTimeParams({"searchMethod": 3, "algoType": "OLD"})


def test_search_before_load_is_value_error():
    idx = nmslib.init(space="l2")
    with pytest.raises(ValueError, match="Must call createIndex or loadIndex"):
        idx.knnQueryBatch([[0.0] * 4], k=1)
    with pytest.raises(ValueError, match="Must call createIndex or loadIndex"):
        idx.saveIndex("/nonexistent/x")
    assert len(idx) == 0


def test_load_missing_file_is_runtime_error(tmp_path):
    idx = nmslib.init(space="l2")
    with pytest.raises(RuntimeError, match="Cannot open file"):
        idx.loadIndex(str(tmp_path / "missing.bin"))


def test_extension_setters_validate_their_arguments():
    # beam width, waves per beam query and precision: accepted ranges and the
    # ValueError of anything outside them (no device needed)
    idx = nmslib.init(space="l2")
    for w in (0, 1, 16, 64):
        idx.setBeam(w)
    for bad in (-1, 65):
        with pytest.raises(ValueError, match="beam must be"):
            idx.setBeam(bad)
    for n in (0, 1, 2, 8):
        idx.setBeamWaves(n)
    for bad in (-1, 3, 4):
        with pytest.raises(ValueError, match="beam waves must be"):
            idx.setBeamWaves(bad)
    for v in (0, 1, 2):
        idx.setScreen(v)
    for bad in (-1, 3):
        with pytest.raises(ValueError, match="screen must be"):
            idx.setScreen(bad)
    # the screen's counters read through their own entry point, zero before a search
    st = idx.lastStats()
    assert st["screened"] == 0 and st["screen_pass"] == 0
    for v in (0, 1, 2):
        idx.setVisitedMarks(v)
    for bad in (-1, 3):
        with pytest.raises(ValueError, match="visited marks must be"):
            idx.setVisitedMarks(bad)
    idx.setPrecision("bf16", 2)
    idx.setPrecision("fp32", 0)
    with pytest.raises(ValueError):
        idx.setPrecision("bf16", -1)


def test_load_rejects_upper_neighbour_without_that_level(tmp_path):
    # the greedy descent reads a level-L neighbour's own list at level L
    # (hnsw_distfunc_opt.cc:168-198); a valid HNSW links a node at level L
    # only to nodes that reach level L, so a file that breaks this is refused
    # at load instead of sending the device scan past the end of the lists
    import struct

    import numpy as np

    from nmslib_amd import indexpack
    from tests import fixtures as F

    path = F.index_path("tiny_l2")
    ix = indexpack.read_index(path)
    h, levels = ix["header"], ix["levels"]
    data = bytearray(open(path, "rb").read())
    # walk the per-node upper-level blocks after the level-0 block
    pos = indexpack.HEADER.size + h["n"] * h["elem"]
    target = None
    for i in range(h["n"]):
        (sm,) = struct.unpack_from("<I", data, pos)
        pos += 4
        if sm and target is None:
            cnt = struct.unpack_from("<i", data, pos)[0]  # level-1 list of node i
            if cnt > 0:
                target = pos + 4  # its first neighbour
        pos += sm
    assert target is not None
    level0_only = int(np.nonzero(levels == 0)[0][0])
    struct.pack_into("<i", data, target, level0_only)
    bad = tmp_path / "bad_level.hnsw"
    bad.write_bytes(bytes(data))
    idx = nmslib.init(space="l2")
    with pytest.raises(RuntimeError, match="has no list at level 1"):
        idx.loadIndex(str(bad))
    idx.loadIndex(path)  # the untouched file loads


def test_beam_gather_setter_validates():
    idx = nmslib.init(space="l2")
    for g in (0, 1, 2):
        idx.setBeamGather(g)
    with pytest.raises(ValueError, match="beam gather"):
        idx.setBeamGather(3)

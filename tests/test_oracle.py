"""Pins the CPU oracle (oracle/) before it is trusted as the checker:
reference golden vectors, the reference's own distance KATs, and what the
mathematics fixes (FP64 closed forms, brute force on tiny graphs)."""
import json
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests import fixtures as F


@pytest.mark.parametrize("name", F.NAMES)
def test_oracle_matches_reference_knnquerybatch(name):
    ix = O.OracleIndex(F.index_path(name))
    res = F.results(name)
    q = res["queries"]
    for ef, k, algo in F.runs(name):
        t = F.tag(ef, k, algo)
        ids, dists, counts = ix.search(q, k, ef, algo, num_threads=4)
        np.testing.assert_array_equal(counts, res[t + "_counts"], err_msg=t)
        ci, cd = F.canon(ids, dists)
        ri, rd = F.canon(res[t + "_ids"], res[t + "_dists"])
        np.testing.assert_array_equal(ci, ri, err_msg=t)
        np.testing.assert_array_equal(cd.view(np.uint32), rd.view(np.uint32), err_msg=t)
        # apart from exactly tied distances the order is the reference's own
        untied = np.ones(ids.shape, bool)
        eq = dists[:, 1:] == dists[:, :-1]
        untied[:, 1:] &= ~eq
        untied[:, :-1] &= ~eq
        np.testing.assert_array_equal(ids[untied], res[t + "_ids"][untied], err_msg=t)


def test_kat_distances():
    kat = json.load(open(os.path.join(F.GOLDEN, "kat_distances.json")))
    v = np.array(kat["vectors"], np.float32)
    l2 = np.array(kat["l2_sqrt"])
    nd = np.array(kat["negdotprod"])
    l1 = np.array(kat["l1"])
    linf = np.array(kat["linf"])
    for i in range(8):
        for j in range(8):
            d2 = O.dist(O.L2SQR, v[i], v[j])  # dim 5: the any-dim kernel
            assert abs(np.sqrt(d2) - l2[i, j]) <= 1e-5
            assert abs(O.dist(O.NEGDOT, v[i], v[j]) - nd[i, j]) <= 1e-5
            assert abs(O.dist(O.L1, v[i], v[j]) - l1[i, j]) <= 1e-5
            assert abs(O.dist(O.LINF, v[i], v[j]) - linf[i, j]) <= 1e-5


@pytest.mark.parametrize("dim", [1, 3, 4, 5, 15, 16, 17, 64, 67, 100, 128])
def test_l1_linf_against_fp64(dim):
    rng = np.random.default_rng(100 + dim)
    for _ in range(20):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        exact = float(np.abs(a.astype(np.float64) - b).sum())
        assert abs(O.dist(O.L1, a, b) - exact) <= 1e-5 * exact
        # |a - b| of floats is exact, so the largest one is too
        assert O.dist(O.LINF, a, b) == np.abs(a - b).max()


def test_l1_summation_order_is_the_sse_one():
    # L1NormSIMD (distcomp_lp.cc:191-247): four float lanes over groups of 4,
    # lanes summed in float, the tail added in double; a plain float loop
    # rounds differently on some of these inputs
    rng = np.random.default_rng(7)
    differs = 0
    for _ in range(200):
        dim = 67
        a = (rng.standard_normal(dim) * 10 ** rng.uniform(-3, 3, dim)).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        t = np.abs(a - b)
        lanes = np.zeros(4, np.float32)
        for g in range(dim // 4):
            lanes = (lanes + t[4 * g:4 * g + 4]).astype(np.float32)
        s = np.float32(np.float32(np.float32(lanes[0] + lanes[1]) + lanes[2]) + lanes[3])
        res = float(s)
        for c in range(4 * (dim // 4), dim):
            res += float(t[c])
        assert O.dist(O.L1, a, b) == np.float32(res)
        seq = np.float32(0)
        for v in t:
            seq = np.float32(seq + v)
        differs += seq != np.float32(res)
    assert differs > 0


@pytest.mark.parametrize("dim", list(range(16, 257, 16)))
def test_l2sqr16_against_fp64(dim):
    # test_distfunc.cc:449-483 (TestL2Sqr16ExtAVXAgree): relative 1e-5
    rng = np.random.default_rng(dim)
    for _ in range(20):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        exact = float(((a.astype(np.float64) - b) ** 2).sum())
        got = O.dist(O.L2SQR16, a, b)
        assert abs(got - exact) <= 1e-5 * exact
        assert O.dist(O.L2SQR, a, b) == pytest.approx(exact, rel=1e-5)


@pytest.mark.parametrize("dim", [1, 3, 4, 5, 15, 17, 96, 100, 127])
def test_dot_kernels_against_fp64(dim):
    rng = np.random.default_rng(1000 + dim)
    for _ in range(20):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        exact = float(a.astype(np.float64) @ b)
        scale = float(np.abs(a.astype(np.float64) * b).sum())
        assert abs(O.dist(O.NEGDOT, a, b) + exact) <= 1e-6 * scale
        an, bn = O.normalize(a), O.normalize(b)
        c = float(an.astype(np.float64) @ bn)
        assert abs(O.dist(O.NORMCOS, an, bn) - max(0.0, 1 - min(1.0, max(-1.0, c)))) <= 2e-6


def test_l2sqr16_summation_order_is_the_avx_one():
    # Values chosen so that a different association of the 8 lane sums gives
    # a different float: lane 0 carries 2^24, the others carry 1 each.
    a = np.zeros(16, np.float32)
    b = np.zeros(16, np.float32)
    a[0] = 4096.0  # (4096)^2 = 2^24 in lane 0
    a[1:8] = 1.0
    # left-to-right: 2^24 + 1 + ... rounds each +1 away -> 2^24
    assert O.dist(O.L2SQR16, a, b) == np.float32(2 ** 24)
    # a pairwise sum would give 2^24 + 8 (exactly representable)
    assert np.float32(2 ** 24) != np.float32(2 ** 24 + 8)


def test_normalize_matches_reference_float_rounding():
    rng = np.random.default_rng(5)
    v = rng.standard_normal(100).astype(np.float32)
    from nmslib_amd.indexpack import normalize_like_reference
    np.testing.assert_array_equal(O.normalize(v), normalize_like_reference(v[None])[0])
    z = np.zeros(7, np.float32)
    np.testing.assert_array_equal(O.normalize(z), z)  # zero sum is left alone


@pytest.mark.parametrize("name", F.NAMES)
def test_oracle_brute_force_vs_reference_seq_search(name):
    m = F.META[name]
    res = F.results(name)
    x = F.base_vectors(name)
    q = res["queries"]
    k = m["bf_k"]
    ids, d64 = O.brute_force(x, q, k, m["space"])
    ref_ids, ref_d = res["bf_ids"], res["bf_dists"]
    if m["space"] == "l2":
        d64 = np.sqrt(d64)  # seq_search l2 is the square root (space_lp.h:57-58)
    # distances agree within the reference's 1e-5; where the ids differ, the
    # reference's id must sit at the same exact (FP64) distance up to float
    # rounding of the reference's single-precision ranking
    np.testing.assert_allclose(ref_d, d64, rtol=1e-5, atol=1e-6)
    x64 = x.astype(np.float64)
    for qi in range(len(q)):
        for j in np.nonzero(ids[qi] != ref_ids[qi])[0]:
            v = x64[ref_ids[qi, j]]
            qq = q[qi].astype(np.float64)
            if m["space"] == "l2":
                d = np.sqrt(((qq - v) ** 2).sum())
            elif m["space"] == "l1":
                d = np.abs(qq - v).sum()
            elif m["space"] == "linf":
                d = np.abs(qq - v).max()
            elif m["space"] == "negdotprod":
                d = -(qq @ v)
            else:
                den = np.linalg.norm(qq) * np.linalg.norm(v)
                d = 1 - (qq @ v) / den if den > 0 else 1.0
            assert abs(d - d64[qi, j]) <= 1e-5 * max(1.0, abs(d)), (qi, j)


def test_tiny_graph_with_ef_at_least_n_is_exact():
    name = "tiny_l2"
    ix = O.OracleIndex(F.index_path(name))
    x = F.base_vectors(name)
    q = F.results(name)["queries"]
    ids, dists, counts = ix.search(q, 10, 200)
    gt, gd = O.brute_force(x, q, 10, "l2")
    np.testing.assert_array_equal(ids, gt)
    np.testing.assert_allclose(dists, gd, rtol=1e-6)


def test_recall_of_uniform10k_grows_with_ef():
    # Uniform [0,1)^128 has a high intrinsic dimension, so recall stays below
    # the final128 ranges of test_integr.cc:86-89; what holds is that the
    # reference's recall grows with ef and that ef >= N would be exact.
    name = "u10k_l2"
    res = F.results(name)
    x = F.base_vectors(name)
    gt, gd = O.brute_force(x, res["queries"], 100, "l2")
    rec = [F.recall_tie_aware(res[f"ef{ef}_k10_hybrid_ids"], gt, gd, 10) for ef in (10, 32, 100, 200)]
    assert all(a < b for a, b in zip(rec, rec[1:])), rec
    assert rec[-1] > 0.9


@pytest.mark.parametrize("name", ["u128_l2", "dec100_cos", "gmm96_nd"])
@pytest.mark.parametrize("k,ef", [(10, 64), (100, 32), (10, 300)])
def test_beam_model_with_one_expansion_is_v1merge(name, k, ef):
    # With one expansion per step and no tied distances, the beam model is
    # SearchV1Merge step for step: same results and the same work counters.
    ix = O.OracleIndex(F.index_path(name))
    q = F.results(name)["queries"][1:]  # query 0 of the cosine set is all zeros (all ties)
    a = ix.search(q, k, ef, "v1merge", with_stats=True)
    b = ix.search_beam(q, k, ef, 1, with_stats=True)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    assert a[3] == b[3]


@pytest.mark.parametrize("beam", [1, 4, 16])
def test_beam_model_tiny_graph_with_ef_at_least_n_is_exact(beam):
    name = "tiny_l2"
    ix = O.OracleIndex(F.index_path(name))
    x = F.base_vectors(name)
    q = F.results(name)["queries"]
    ids, dists, _ = ix.search_beam(q, 10, 200, beam)
    gt, gd = O.brute_force(x, q, 10, "l2")
    np.testing.assert_array_equal(ids, gt)
    np.testing.assert_allclose(dists, gd, rtol=1e-6)


@pytest.mark.parametrize("beam", [2, 8, 16])
def test_beam_model_keeps_quality(beam):
    name = "u128_l2"
    ix = O.OracleIndex(F.index_path(name))
    x = F.base_vectors(name)
    q = F.results(name)["queries"]
    ex_ids, _, _ = ix.search(q, 10, 64, "v1merge")
    ids, dists, counts, st = ix.search_beam(q, 10, 64, beam, with_stats=True)
    assert np.all(counts == 10)
    assert np.all(np.diff(dists, axis=1) >= 0)
    for i in range(len(q)):
        assert len(set(ids[i].tolist())) == 10
        for j in (0, 9):
            assert dists[i, j] == O.dist(ix.func, q[i], x[ids[i, j]])
    gt, gd = O.brute_force(x, q, 100, "l2")
    r_ex = F.recall_tie_aware(ex_ids, gt, gd, 10)
    r_b = F.recall_tie_aware(ids, gt, gd, 10)
    assert r_b >= r_ex - 0.01, (r_b, r_ex)
    assert st["expansions"] >= 64 * len(q)


def test_stats_counts_are_consistent():
    ix = O.OracleIndex(F.index_path("u128_l2"))
    q = F.results("u128_l2")["queries"]
    _, _, _, st = ix.search(q, 10, 64, with_stats=True)
    assert st["expansions"] >= 64 * len(q)  # V1Merge expands at least ef items
    assert st["dist_level0"] <= st["list_ints_read"]
    assert st["dist_upper"] >= len(q)


def _corrupt(src, dst, offset, payload):
    b = bytearray(open(src, "rb").read())
    b[offset:offset + len(payload)] = payload
    open(dst, "wb").write(bytes(b))


def test_loader_errors(tmp_path):
    src = F.index_path("tiny_l2")
    bad = str(tmp_path / "bad.hnsw")
    _corrupt(src, bad, 56, struct.pack("<i", 9))
    with pytest.raises(RuntimeError, match="Unknown distance function code: 9"):
        O.OracleIndex(bad)
    data = open(src, "rb").read()
    open(bad, "wb").write(data[:len(data) // 2])
    with pytest.raises(RuntimeError):
        O.OracleIndex(bad)
    with pytest.raises(RuntimeError, match="Cannot open file"):
        O.OracleIndex(str(tmp_path / "missing.hnsw"))

"""CPU replay of the two-wave beam kernel's command protocol
(hnsw_search_kernel_pair, hnsw_search_impl.h: PairLink::post, beam_helper,
walk_pair, sort_merge) over the beam model's real steps on the 10,000-point
l2 fixture -- the index whose two-wave GPU cases stay skipped -- for every
(beam width, k, efSearch) case of the GPU parity test.

For each step of each sampled query, the first wave's commands are generated
as the kernel posts them (distances when the second wave takes candidates,
one rank and optionally one walk per merge pass of at most 256 items, each
followed by an empty command; with the deferred rebuild the second wave takes
every candidate and a single-pass merge posts no walk, the first wave
rebuilding alone during the next step's distances), the second wave consumes
them as beam_helper does, and the replay checks that:
  * both waves pass the same number of workgroup barriers, step by step;
  * every LDS index the step uses stays inside its region, and the regions
    laid out by carve() end inside the block's LDS allocation (padded, in
    two-wave mode, past a quarter of the CU's LDS so a fourth block never
    fits): the sorted array, the candidate/distance region (ranks of up to
    256 items included), the items region, the merge records, the claimed
    picks, the command words and, deferred, the second items buffer."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import fixtures as F

LDS_PER_CU = 160 * 1024
MAX_STEP_ENTRIES = 512
WIDE_NV = {"fp32": 2, "bf16": 5, "fp8": 8}


def layout(cap, dstride, ent):
    """carve()'s offsets (bytes), lds_wave_bytes(), and pair_lds(): the padded
    launch size and whether the deferred rebuild's second items buffer fits
    without losing a query per CU."""
    regions = {}
    off = 0
    for name, size in [("A", ((cap + 1) & ~1) * 8), ("qs", 2 * dstride * 4), ("X", 2 * ent * 4),
                       ("Y", 2 * ent * 4), ("meta", (cap // 64 + 2) * 16), ("pfa", 24 * 8),
                       ("pfq", 16 * 8), ("pfpos", 24 * 4), ("ctl", 24 * 4), ("Z", 8 * ent)]:
        regions[name] = (off, off + size)
        off += size
    b = ((cap + 1) & ~1) * 8 + dstride * 8 + 16 * ent + (cap // 64 + 2) * 16 + 416 + 96
    lds = (b + 15) // 16 * 16
    floor = LDS_PER_CU // 4 + 16
    plain, with2 = max(lds, floor), max(lds + 8 * ent, floor)
    defer = LDS_PER_CU // with2 == LDS_PER_CU // plain
    return regions, off, lds, (with2 if defer else plain), defer


def replay(trace, W, E, C, ef, storage, regions, defer):
    """Barrier counts of both waves per step, with the step's LDS bounds checked."""
    P0 = 0 if storage == "fp32" else 32          # first wave's candidates per round
    P1 = 32 * WIDE_NV[storage]                   # second wave's (the wide path)
    claim = C == ef
    out = []
    n_expect = 1
    for (n, nsel, ncand, m, p0a, p0b) in trace:
        assert n == n_expect, "array size differs from the previous step's merge"
        assert 1 <= nsel <= W <= 16            # picks / pfq / pfa indices
        assert ncand <= E                      # cand[] and cdist[] in region X
        assert m <= ncand
        cmds, first = [], 0

        def post(cmd, *p):
            nonlocal first
            cmds.append((cmd,) + p)
            first += 1

        r = -(-ncand // (P0 + P1))
        h0 = 0 if defer else max(0, ncand - P1 * r)
        if h0 < ncand:
            post("dist", ncand, h0)
            post("nop")
            # the second wave's admitted items at items[h0 .. h0 + m1): in Y
            assert ncand - h0 <= E
        single = m <= 256 and claim
        passes = [(0, m)] if single else [(b, min(256, m - b)) for b in range(0, m, 256)]
        assert len(passes) <= 2
        np_ = n
        for i, (b, cnt) in enumerate(passes):
            if cnt == 0:
                continue
            nout = min(C, np_ + cnt)
            p0 = (p0a, p0b)[i]
            assert 0 <= p0 <= np_
            assert cnt <= 256 and cnt <= 2 * E  # the ranks come back in cand[0..cnt)
            post("rank", np_, cnt, nout, b)
            post("nop")
            if p0 < nout:
                clo, chi = p0 >> 6, (nout - 1) >> 6
                assert chi - clo < C // 64 + 2  # merge records
                assert (chi << 6) + 63 < (regions["A"][1] - regions["A"][0]) // 8 + 64
                if defer and single:
                    # the first wave walks alone in the next step (walk_all):
                    # no barrier; its items buffer stays untouched until then
                    np_ = nout
                    continue
                post("walk", clo, chi, nout, b)
                nr = (chi - clo + 16) >> 4
                first += nr  # walk_pair's barriers on the first wave
                post("nop")
            np_ = nout
        n_expect = np_
        # the second wave: one barrier per command at the top of its loop,
        # plus walk_pair's rounds derived from the command's own words
        helper = 0
        for c in cmds:
            helper += 1
            if c[0] == "walk":
                helper += (c[2] - c[1] + 16) >> 4
        out.append((first, helper))
    return out


@pytest.mark.parametrize("defer", [False, True])
@pytest.mark.parametrize("storage", ["fp32", "bf16"])
@pytest.mark.parametrize("beam", [1, 2, 5, 8, 16])
def test_pair_protocol_replay_on_the_10k_fixture(beam, storage, defer):
    name = "u10k_l2"
    path = F.index_path(name)
    o = O.OracleIndex(path)
    q = F.results(name)["queries"]
    dstride = (o.dim + 3) // 4 * 4
    W = max(1, min(beam, MAX_STEP_ENTRIES // o.maxM0))
    E = max(128, (W * o.maxM0 + 63) // 64 * 64)
    steps = 0
    for k, ef in [(10, 64), (100, 32), (10, 300), (1, 700)]:
        C = max(ef, k)
        regions, end, lds, lds_launch, fits = layout(C, dstride, E)
        if defer and not fits:
            continue
        # carve() fits the allocation (the second items buffer only when deferred)
        assert (end if defer else regions["Z"][0]) <= lds_launch
        assert lds <= lds_launch
        assert 4 * lds_launch > LDS_PER_CU >= 3 * lds_launch  # three blocks per CU, never four
        for i in range(0, len(q), 25):
            tr = o.beam_trace(q[i], k, ef, W)
            counts = replay(tr, W, E, C, ef, storage, regions, defer)
            for s, (a, b) in enumerate(counts):
                assert a == b, f"query {i} k={k} ef={ef} step {s}: barriers {a} vs {b}"
            steps += len(tr)
    assert steps > 1000

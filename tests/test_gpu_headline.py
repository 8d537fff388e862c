"""The benchmark's kernel at the benchmark's shape, against the oracle.

The batched hexagon CFTP (lozenge hexagon a = b = c = 300, N = 609) runs
its chain pairs on tile shape 7 (k_pair_lozenge: both chains in one LDS
word per cell, 96 x 32 tiles, halo 9 rows and 3 columns, eight waves per
SIMD) with far more workgroups than the chip holds at once: 64 tiles per
sample, 16 samples = 1024 workgroups against 2 per CU.  Here both
chains of every sample after exactly three doubling rounds (look-back
1024 + 1024 + 2048 = 4096 steps in the last round, 7168 per sample in all)
come back through rt_lozenge_cftp_batch_rounds and are compared bit for
bit with the oracle's segmented walks over that round's segment list
(LozengeTiler.cpp:27-159 per segment, seeds newest first as the CFTP
driver prepends them, AztecDiamondCftpEx.cpp:57-76), from MaxTiling and
MinTiling.  The other pair shapes (64-row pairs, 64-row split pairs, at
halo 6) must give the same chains.
"""
import os

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu

SIDE = 300
T0 = 1024
ROUNDS = 3
MASTERS = list(range(1, 17))


def _say(*a):
    print("[case]", *a, flush=True)


def round_segments(master, rounds, t0):
    """Segment lengths and seeds of doubling round `rounds`, newest first."""
    seeds = [int(s) for s in o.mt19937(master, rounds)]
    steps = [t0]
    for _ in range(rounds - 1):
        steps.insert(0, sum(steps))
    return steps, seeds[::-1]


@pytest.fixture(scope="module")
def hexagon():
    d = rt.LozengeTiler.AlmostHexDomain(SIDE, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 300)
    return tmin, tmax, P


def test_round_segments_follow_the_driver():
    steps, seeds = round_segments(1, 3, 1024)
    assert steps == [2048, 1024, 1024]
    assert seeds[-1] == int(o.mt19937(1, 1)[0])


def test_headline_pair_kernel_matches_oracle_at_hexagon_300(gpu_ctx, hexagon):
    tmin, tmax, P = hexagon
    gpu_ctx.set_tinymt_params(P)
    runs = {}
    try:
        for shape in (7, 2, 5):
            gpu_ctx.set_tile_shape(shape)
            _say("hexagon 300 cftp pairs, 16 samples, 3 rounds, tile shape", shape)
            tops, bots, stats = gpu_ctx.lozenge_cftp_batch_rounds(
                tmax, tmin, MASTERS, rt.SCHED_DOUBLING, T0, ROUNDS)
            tm = gpu_ctx.last_timing()
            assert tm["shape"] == (shape if shape != 5 else 2)
            assert all(s["rounds"] == ROUNDS for s in stats)
            runs[shape] = (tops, bots)
    finally:
        gpu_ctx.set_tile_shape(0)
    tops, bots = runs[7]
    for shape in (2, 5):
        assert np.array_equal(runs[shape][0], tops), shape
        assert np.array_equal(runs[shape][1], bots), shape
    # the chains moved and did not all coalesce in 4096 steps
    assert not np.array_equal(tops[0], tmax) and not np.array_equal(bots[0], tmin)
    o.set_threads(min(os.cpu_count() or 1, 16))
    try:
        for k in (0, 7, 15):
            steps, seeds = round_segments(MASTERS[k], ROUNDS, T0)
            _say("oracle walks for master", MASTERS[k], steps)
            want_top = o.lozenge_walk_segments(tmax, steps, seeds, P)
            assert np.array_equal(tops[k], want_top), MASTERS[k]
            want_bot = o.lozenge_walk_segments(tmin, steps, seeds, P)
            assert np.array_equal(bots[k], want_bot), MASTERS[k]
    finally:
        o.set_threads(1)


def test_stream_split_never_changes_results(gpu_ctx, hexagon):
    """The batch's units split over 1, 2 or 3 streams (the default is 2):
    the same chains after two doubling rounds of 9 hexagon-300 samples
    (576 workgroups, enough to split; odd, so the parts differ in size)."""
    tmin, tmax, P = hexagon
    gpu_ctx.set_tinymt_params(P)
    runs = []
    try:
        for n in (1, 2, 3):
            _say("hexagon 300 cftp pairs, 9 samples, 2 rounds, streams", n)
            gpu_ctx.set_streams(n)
            tops, bots, _ = gpu_ctx.lozenge_cftp_batch_rounds(
                tmax, tmin, [3, 1, 4, 1, 5, 9, 2, 6, 5], rt.SCHED_DOUBLING, T0, 2)
            runs.append((tops, bots))
    finally:
        gpu_ctx.set_streams(2)
    for tops, bots in runs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(tops, runs[0][0]))
        assert all(np.array_equal(a, b) for a, b in zip(bots, runs[0][1]))
    # a master seed twice (1 at 1 and 3, 5 at 4 and 8): the same chains
    assert np.array_equal(runs[0][0][1], runs[0][0][3])
    assert np.array_equal(runs[0][1][4], runs[0][1][8])
    with pytest.raises(rt.RTError):
        gpu_ctx.set_streams(0)
    with pytest.raises(rt.RTError):
        gpu_ctx.set_streams(5)


def test_rounds_entry_point_agrees_with_cftp_batch(gpu_ctx):
    """On a small hexagon: where the CFTP batch says a sample coalesced in
    round r, the fixed-rounds call at r returns two equal chains equal to
    that sample."""
    d = rt.LozengeTiler.AlmostHexDomain(9, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 301)
    gpu_ctx.set_tinymt_params(P)
    masters = [3, 4, 5]
    _say("hexagon 9 cftp batch vs fixed rounds")
    outs, stats = gpu_ctx.lozenge_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, 64, 20)
    for k, m in enumerate(masters):
        r = stats[k]["rounds"]
        tops, bots, _ = gpu_ctx.lozenge_cftp_batch_rounds(tmax, tmin, [m], rt.SCHED_DOUBLING,
                                                          64, r)
        assert np.array_equal(tops[0], outs[k]) and np.array_equal(bots[0], outs[k])
        if r > 1:
            tops, bots, _ = gpu_ctx.lozenge_cftp_batch_rounds(tmax, tmin, [m],
                                                              rt.SCHED_DOUBLING, 64, r - 1)
            assert not np.array_equal(tops[0], bots[0])
    with pytest.raises(rt.RTError):
        gpu_ctx.lozenge_cftp_batch_rounds(tmax, tmin, [1], rt.SCHED_DOUBLING, 64, 0)


@pytest.mark.parametrize("halo", [9, 6, 5])
def test_lozenge_walk_on_shape_7_matches_oracle(gpu_ctx, halo):
    """A one-chain lozenge Walk (LozengeTiler.cpp:27-159) forced onto tile
    shape 7: at halos divisible by 3 k_pair_lozenge runs it with the chain
    in the low halves of its words, otherwise k_walk_o8<1>; the hexagon of
    side 100 (201 rows, 67 plane columns: several 96 x 32 tiles)."""
    d = rt.LozengeTiler.AlmostHexDomain(100, 0)
    tmin = rt.LozengeTiler.MinTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 101)
    gpu_ctx.set_tinymt_params(P)
    want = o.lozenge_walk(tmin, 500, 3, P)
    steps, seeds = [200, 100], [5, 9]
    want_seg = o.lozenge_walk_segments(tmin, steps, seeds, P)
    try:
        gpu_ctx.set_halo(halo)
        gpu_ctx.set_tile_shape(7)
        _say("lozenge walk hexagon 100, shape 7, halo", halo)
        assert np.array_equal(gpu_ctx.lozenge_walk(tmin, 500, 3), want)
        assert gpu_ctx.last_timing()["shape"] == 7
        assert np.array_equal(gpu_ctx.lozenge_walk_segments(tmin, steps, seeds), want_seg)
    finally:
        gpu_ctx.set_halo(0)
        gpu_ctx.set_tile_shape(0)

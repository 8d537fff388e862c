"""Batched device paths: every member of a batch equals the same work run
alone, and the oracle."""
import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_lozenge_cftp_batch_equals_singles(gpu_ctx):
    d = rt.LozengeTiler.AlmostHexDomain(6, 1)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 21)
    gpu_ctx.set_tinymt_params(P)
    masters = [1, 2, 3, 4, 5, 17, 345903]
    for sched in (rt.SCHED_DOUBLING, rt.SCHED_REFERENCE):
        _say("lozenge cftp batch", sched)
        outs, stats = gpu_ctx.lozenge_cftp_batch(tmax, tmin, masters, sched, 64, 30)
        for k, m in enumerate(masters):
            want, ws = o.lozenge_cftp(tmax, tmin, m, P, sched, 64, 30)
            assert np.array_equal(outs[k], want), (sched, m)
            assert stats[k]["rounds"] == ws["rounds"]


def test_domino_cftp_batch_equals_singles(gpu_ctx):
    d = rt.DominoTiler.AztecDiamond(6)
    tmin, tmax = rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, 22)
    gpu_ctx.set_tinymt_params(P)
    masters = [345903, 1, 2, 3, 9]
    for rule, a, b in [(rt.DOMINO_HEATBATH, 1.0, 1.0), (rt.DOMINO_HEATBATH2P, 1.0, 0.5)]:
        _say("domino cftp batch", rule)
        outs, stats = gpu_ctx.domino_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, 128, 30,
                                                rule, a, b)
        for k, m in enumerate(masters):
            single, _ = gpu_ctx.domino_cftp(tmax, tmin, m, rt.SCHED_DOUBLING, 128, 30, rule, a, b)
            want, _ = o.domino_cftp(tmax, tmin, m, P, rt.SCHED_DOUBLING, 128, 30, rule, a, b)
            assert np.array_equal(outs[k], single)
            assert np.array_equal(outs[k], want)


def test_domino_cftp_rounds_match_oracle_segments(gpu_ctx):
    """rt_domino_cftp_batch_rounds: after r doubling rounds both chains of
    every sample equal the oracle's segmented walks from MaxTiling and
    MinTiling (DominoTiler.cpp:125-231, segment lengths 2^(r-2) t0, ...,
    t0, t0 and seeds newest first, AztecDiamondCftpEx.cpp:57-76); at the
    round where the CFTP batch reports coalescence the chains are equal and
    equal to its sample."""
    d = rt.DominoTiler.AztecDiamond(10)
    tmin, tmax = rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, 23)
    gpu_ctx.set_tinymt_params(P)
    masters = [1, 2, 7]
    rule = (rt.DOMINO_HEATBATH2P, 1.0, 0.5)
    _say("domino cftp rounds vs oracle segments")
    tops, bots, stats = gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, masters, rt.SCHED_DOUBLING,
                                                         64, 3, *rule)
    assert all(s["rounds"] == 3 for s in stats)
    for k, m in enumerate(masters):
        seeds = [int(x) for x in o.mt19937(m, 3)][::-1]
        steps = [128, 64, 64]
        assert np.array_equal(tops[k], o.domino_walk_segments(tmax, steps, seeds, P, *rule)), m
        assert np.array_equal(bots[k], o.domino_walk_segments(tmin, steps, seeds, P, *rule)), m
    outs, cs = gpu_ctx.domino_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, 64, 30, *rule)
    for k, m in enumerate(masters):
        r = cs[k]["rounds"]
        t, b, _ = gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, [m], rt.SCHED_DOUBLING, 64, r,
                                                   *rule)
        assert np.array_equal(t[0], outs[k]) and np.array_equal(b[0], outs[k]), m
    with pytest.raises(rt.RTError):
        gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, [1], rt.SCHED_DOUBLING, 64, 2,
                                         rt.DOMINO_METRO2P, 1.0, 0.5)


@pytest.mark.parametrize("use_graph", [False, True])
def test_domino_walk_batch_uneven(gpu_ctx, use_graph):
    orders = [3, 17, 40, 9, 64, 25]
    tilings = []
    for n in orders:
        d = rt.DominoTiler.AztecDiamond(n)
        tilings.append(rt.DominoTiler.MinTiling(d))
    tilings.append(rt.DominoTiler.MaxTiling(rt.DominoTiler.Rectangle(30, 48)))
    big = max(rt.side(t) for t in tilings)
    P = rt.synthetic_params(big * big // 2, 23)
    gpu_ctx.set_tinymt_params(P)
    seeds = [11 * k + 1 for k in range(len(tilings))]
    _say("domino walk batch graph=", use_graph)
    outs = gpu_ctx.domino_walk_batch(tilings, 1500, seeds, rt.DOMINO_TOGGLE08, use_graph=use_graph)
    for t, s, got in zip(tilings, seeds, outs):
        want = o.domino_walk(t, 1500, s, P, rt.DOMINO_TOGGLE08)
        assert np.array_equal(got, want)


def test_domino_walk_batch_rect_unions(gpu_ctx):
    from randomtilingspublic_amd import workloads as W
    batch = W.uneven_domino_batch(5, 10, order=(3, 40), side=(4, 40), max_extent=64)
    tilings = [b["tiling"] for b in batch]
    big = max(b["N"] for b in batch)
    P = rt.synthetic_params(big * big // 2, 24)
    gpu_ctx.set_tinymt_params(P)
    seeds = [7 * k + 3 for k in range(len(tilings))]
    # the 0.8 toggle: the autotuner picks its tile shape (heat-bath walks go
    # to the packed kernel directly, checked below)
    want = [o.domino_walk(t, 1800, s, P, rt.DOMINO_TOGGLE08) for t, s in zip(tilings, seeds)]
    gpu_ctx.set_autotune(2)
    try:
        for use_graph in (False, True):
            _say("domino walk batch rect unions graph=", use_graph)
            outs = gpu_ctx.domino_walk_batch(tilings, 1800, seeds, rt.DOMINO_TOGGLE08,
                                             use_graph=use_graph)
            tm = gpu_ctx.last_timing()
            # first call tunes (1800 steps: ~150 launches), the second reuses it
            assert tm["shape_source"] == (2 if use_graph else 1), tm
            assert tm["shape"] in (1, 2, 3, 7)
            for got, w in zip(outs, want):
                assert np.array_equal(got, w)
        _say("domino walk batch rect unions, heat bath")
        outs = gpu_ctx.domino_walk_batch(tilings, 1800, seeds, rt.DOMINO_HEATBATH)
        tm = gpu_ctx.last_timing()
        assert tm["shape"] == 7 and tm["shape_source"] == 0, tm
        for got, t, s in zip(outs, tilings, seeds):
            assert np.array_equal(got, o.domino_walk(t, 1800, s, P, rt.DOMINO_HEATBATH))
    finally:
        gpu_ctx.set_autotune(1)


def test_tile_shape_choice_never_changes_results(gpu_ctx):
    d = rt.LozengeTiler.AlmostHexDomain(9, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 25)
    gpu_ctx.set_tinymt_params(P)
    masters = [3, 8, 13]
    ref = None
    try:
        for mode, shape in [(0, 0), (2, 0), (1, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6),
                            (0, 7)]:
            _say("lozenge cftp batch autotune", mode, "shape", shape)
            gpu_ctx.set_autotune(mode)
            gpu_ctx.set_tile_shape(shape)
            outs, stats = gpu_ctx.lozenge_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, 2048,
                                                     30)
            tm = gpu_ctx.last_timing()
            if shape:
                assert tm["shape"] == (shape - 3 if 3 < shape < 7 else shape)
                assert tm["shape_source"] == 0 and tm["split"] == (3 < shape < 7)
            if mode == 2:
                assert tm["shape_source"] == 1, tm
            if ref is None:
                ref = outs
                for k, m in enumerate(masters):
                    want, _ = o.lozenge_cftp(tmax, tmin, m, P, rt.SCHED_DOUBLING, 2048, 30)
                    assert np.array_equal(outs[k], want)
            else:
                for a, b in zip(outs, ref):
                    assert np.array_equal(a, b)
    finally:
        gpu_ctx.set_autotune(1)
        gpu_ctx.set_tile_shape(0)


def test_cftp_pair_with_equal_chains_runs_one_chain_exactly(gpu_ctx):
    # top == bottom from the start: every tile of every window takes the
    # one-chain path of the pair kernels; the sample is the plain walk's.
    d = rt.LozengeTiler.AlmostHexDomain(10, 2)
    tmin = rt.LozengeTiler.MinTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 26)
    gpu_ctx.set_tinymt_params(P)
    try:
        for shape in (1, 2, 3, 6, 7):
            _say("lozenge cftp equal chains shape", shape)
            gpu_ctx.set_tile_shape(shape)
            outs, stats = gpu_ctx.lozenge_cftp_batch(tmin, tmin, [5, 6], rt.SCHED_DOUBLING, 300, 4)
            for k, m in enumerate((5, 6)):
                want, ws = o.lozenge_cftp(tmin, tmin, m, P, rt.SCHED_DOUBLING, 300, 4)
                assert stats[k]["rounds"] == ws["rounds"] == 1
                assert np.array_equal(outs[k], want)
    finally:
        gpu_ctx.set_tile_shape(0)


def test_walk_batch_with_more_tiles_than_fit_on_the_chip(gpu_ctx):
    # 96 Aztec diamonds of order 100 = 768 tiles of 64 rows: many workgroups
    # start after a neighbouring tile's workgroup has finished its window.
    # Each chain must still equal its lone walk (8 tiles, all resident) and
    # the oracle.
    d = rt.DominoTiler.AztecDiamond(100)
    t = rt.DominoTiler.MinTiling(d)
    N = rt.side(t)
    P = rt.synthetic_params(N * N // 2, 27)
    gpu_ctx.set_tinymt_params(P)
    n = 96
    seeds = [1000 + 17 * k for k in range(n)]
    try:
        for shape in (1, 2):
            gpu_ctx.set_tile_shape(shape)
            _say("walk batch 96 x Aztec 100 shape", shape)
            outs = gpu_ctx.domino_walk_batch([t] * n, 600, seeds, rt.DOMINO_HEATBATH,
                                             use_graph=True)
            for k in range(0, n, 7):
                single = gpu_ctx.domino_walk(t, 600, seeds[k], rt.DOMINO_HEATBATH)
                assert np.array_equal(outs[k], single), k
            for k in (0, n - 1):
                want = o.domino_walk(t, 600, seeds[k], P, rt.DOMINO_HEATBATH)
                assert np.array_equal(outs[k], want), k
    finally:
        gpu_ctx.set_tile_shape(0)


def domino_counts(t):
    t = np.asarray(t)
    return [int(((t >> b) & 1).sum()) for b in range(4)] + [int(((t == 3) | (t == 12)).sum())]


def test_walk_batch_device_observables(gpu_ctx):
    tilings = [rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(n)) for n in (5, 30, 61)]
    tilings.append(rt.DominoTiler.MinTiling(rt.DominoTiler.RectangleUnion(
        [(0, 0, 20, 30), (10, 24, 30, 12)])))
    big = max(rt.side(t) for t in tilings)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(big * big // 2, 28))
    seeds = [5, 6, 7, 8]
    _say("walk batch observables")
    outs, counts = gpu_ctx.domino_walk_batch_observe(tilings, 700, seeds, rt.DOMINO_HEATBATH)
    for t, c in zip(outs, counts):
        assert list(c) == domino_counts(t)
        assert c[0] == c[1] and c[2] == c[3]  # each domino sets one bit at each end
    plain = gpu_ctx.domino_walk_batch(tilings, 700, seeds, rt.DOMINO_HEATBATH)
    for a, b in zip(outs, plain):
        assert np.array_equal(a, b)
    none, again = gpu_ctx.domino_walk_batch_observe(tilings, 700, seeds, rt.DOMINO_HEATBATH,
                                                    use_graph=True, download=False)
    assert none is None and np.array_equal(again, counts)

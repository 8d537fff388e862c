"""The packed domino chain-pair kernel (k_pair_domino, tile shape 7 with an
even halo) against the oracle, at a size where it runs more workgroups than
the chip holds: the Aztec diamond of order 250 (N = 502, 60 tiles of 96 x 32
cells per sample), 12 master seeds, three doubling rounds of the CFTP
driver (AztecDiamondCftpEx.cpp:57-76) returned by
rt_domino_cftp_batch_rounds.  Both chains of three samples are compared bit
for bit with the oracle's segmented walks from MaxTiling and MinTiling
(DominoTiler.cpp:125-231), for the heat-bath rule (dominokernelCFTP.cl) and
the two-periodic one (dominokernelCFTPtwoperiodic.cl, a = 1, b = 0.5); the
byte-cell pair shapes (64-row pairs, split pairs) must give the same
chains.  A smaller case runs every even halo the kernel takes (and odd ones,
which fall back to the byte-cell kernel).
"""
import os

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu

T0 = 64
ROUNDS = 3
RULES = [(rt.DOMINO_HEATBATH, 1.0, 1.0), (rt.DOMINO_HEATBATH2P, 1.0, 0.5)]


def _say(*a):
    print("[case]", *a, flush=True)


def _segments(master, rounds, t0):
    seeds = [int(x) for x in o.mt19937(master, rounds)][::-1]
    steps = [t0]
    for _ in range(rounds - 1):
        steps.insert(0, sum(steps))
    return steps, seeds


@pytest.fixture(scope="module")
def aztec250():
    d = rt.DominoTiler.AztecDiamond(250)
    tmin, tmax = rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, 250)
    return tmin, tmax, P


@pytest.mark.parametrize("rule", RULES, ids=["heatbath", "heatbath2p"])
def test_pair_domino_matches_oracle_at_aztec_250(gpu_ctx, aztec250, rule):
    tmin, tmax, P = aztec250
    masters = list(range(1, 13))
    gpu_ctx.set_tinymt_params(P)
    runs = {}
    try:
        for shape in (7, 2, 5):
            _say("aztec 250 domino cftp pairs, 12 samples, 3 rounds, rule", rule, "shape", shape)
            gpu_ctx.set_tile_shape(shape)
            tops, bots, stats = gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, masters,
                                                                 rt.SCHED_DOUBLING, T0, ROUNDS,
                                                                 *rule)
            assert gpu_ctx.last_timing()["shape"] == (shape if shape != 5 else 2)
            assert all(s["rounds"] == ROUNDS for s in stats)
            runs[shape] = (tops, bots)
    finally:
        gpu_ctx.set_tile_shape(0)
    tops, bots = runs[7]
    for shape in (2, 5):
        assert np.array_equal(runs[shape][0], tops), shape
        assert np.array_equal(runs[shape][1], bots), shape
    assert not np.array_equal(tops[0], tmax) and not np.array_equal(bots[0], tmin)
    o.set_threads(min(os.cpu_count() or 1, 16))
    try:
        for k in (0, 5, 11):
            steps, seeds = _segments(masters[k], ROUNDS, T0)
            _say("oracle walks for master", masters[k], steps)
            assert np.array_equal(tops[k], o.domino_walk_segments(tmax, steps, seeds, P, *rule))
            assert np.array_equal(bots[k], o.domino_walk_segments(tmin, steps, seeds, P, *rule))
    finally:
        o.set_threads(1)


@pytest.mark.parametrize("halo", [2, 4, 6, 8, 12, 5])
def test_pair_domino_halo_never_changes_results(gpu_ctx, halo):
    """Aztec order 70 (142 rows, 71 plane columns: several tiles at every
    halo), 3 samples, both rules (two-periodic also with weights whose
    ratio is no power of two), tile shape 7 against 16 x 4 pairs at the
    default halo; the oracle for one sample."""
    d = rt.DominoTiler.AztecDiamond(70)
    tmin, tmax = rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, 70)
    gpu_ctx.set_tinymt_params(P)
    masters = [4, 9, 16]
    try:
        for rule in RULES + [(rt.DOMINO_HEATBATH2P, 0.9, 1.3)]:
            gpu_ctx.set_halo(0)
            gpu_ctx.set_tile_shape(2)
            want_t, want_b, _ = gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, masters,
                                                                 rt.SCHED_DOUBLING, 128, 3, *rule)
            _say("aztec 70 domino cftp pairs halo", halo, "rule", rule)
            gpu_ctx.set_halo(halo)
            gpu_ctx.set_tile_shape(7)
            t, b, _ = gpu_ctx.domino_cftp_batch_rounds(tmax, tmin, masters, rt.SCHED_DOUBLING,
                                                       128, 3, *rule)
            assert np.array_equal(t, want_t) and np.array_equal(b, want_b)
            steps, seeds = _segments(masters[1], 3, 128)
            assert np.array_equal(t[1], o.domino_walk_segments(tmax, steps, seeds, P, *rule))
    finally:
        gpu_ctx.set_halo(0)
        gpu_ctx.set_tile_shape(0)


WALK_RULES = [(rt.DOMINO_TOGGLE08, 1.0, 1.0), (rt.DOMINO_HEATBATH, 1.0, 1.0),
              (rt.DOMINO_METRO2P, 1.0, 0.5), (rt.DOMINO_HEATBATH2P, 1.0, 0.5),
              (rt.DOMINO_METRO2P, 0.9, 1.3)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_packed_domino_walk_batch_matches_oracle(gpu_ctx, use_graph):
    """One-chain domino walks on tile shape 7 for every rule (heat-bath
    rules on k_pair_domino, the chain in the low halves of the words; the
    others on k_walk_o8<0>): 40 Aztec diamonds of order
    60..160 in one batch (about 700 workgroups of 96 x 32 tiles, more than
    the chip holds), 600 steps, eager (two streams) and as a hipGraph (two
    parallel chains), against 64-row tiles and, for three chains, the
    oracle's Walk (DominoTiler.cpp:18-122)."""
    orders = [60 + (37 * k) % 101 for k in range(40)]
    tilings = [rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(n)) for n in orders]
    big = max(rt.side(t) for t in tilings)
    P = rt.synthetic_params(big * big // 2, 160)
    gpu_ctx.set_tinymt_params(P)
    seeds = [13 * k + 5 for k in range(len(tilings))]
    try:
        for rule in WALK_RULES:
            runs = {}
            for shape in (7, 2):
                _say("domino walk batch, 40 aztec diamonds, rule", rule, "shape", shape,
                     "graph", use_graph)
                gpu_ctx.set_tile_shape(shape)
                runs[shape] = gpu_ctx.domino_walk_batch(tilings, 600, seeds, rule[0], rule[1],
                                                        rule[2], use_graph=use_graph)
            for a, b in zip(runs[7], runs[2]):
                assert np.array_equal(a, b)
            for k in (0, 17, 39):
                assert np.array_equal(runs[7][k],
                                      o.domino_walk(tilings[k], 600, seeds[k], P, *rule)), k
    finally:
        gpu_ctx.set_tile_shape(0)

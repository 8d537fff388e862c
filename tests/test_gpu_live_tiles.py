"""Batched jobs leave out the tiles whose cells stay zero in every tiling a
chain can reach (the cells farther than one row and one column from the
start tilings' nonzero cells; tests/test_live_cells.py checks that fact on
the oracle).  Domains whose bounding square is mostly empty (an L, a cross,
Aztec diamonds), one-chain walks on every kernel family of tile shape 7 and
the 64-row tiles, and lozenge CFTP pairs on a hexagon with a slanted side:
bit for bit against the oracle, with fewer sites computed than the grid
holds."""
import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _domino_sites(t):
    N = rt.side(t)
    return (N - 2) * (N // 2 - 2)


RULES = [(rt.DOMINO_HEATBATH, 1.0, 1.0), (rt.DOMINO_TOGGLE08, 1.0, 1.0),
         (rt.DOMINO_HEATBATH2P, 0.9, 1.3)]


@pytest.mark.parametrize("rule", RULES, ids=["heatbath", "toggle", "heatbath2p"])
def test_walks_on_sparse_domains_match_oracle(gpu_ctx, rule):
    D = rt.DominoTiler
    tilings = [D.MinTiling(D.RectangleUnion([(0, 0, 240, 14), (226, 0, 14, 240)])),
               D.MaxTiling(D.RectangleUnion([(0, 100, 220, 16), (100, 0, 16, 220)])),
               D.MinTiling(D.AztecDiamond(90)), D.MaxTiling(D.AztecDiamond(57))]
    big = max(rt.side(t) for t in tilings)
    P = rt.synthetic_params(big * big // 2, 41)
    gpu_ctx.set_tinymt_params(P)
    seeds = [3, 14, 15, 92]
    steps = 900
    try:
        for shape in (7, 2):
            _say("sparse-domain walk batch, rule", rule, "shape", shape)
            gpu_ctx.set_tile_shape(shape)
            outs = gpu_ctx.domino_walk_batch(tilings, steps, seeds, *rule)
            tm = gpu_ctx.last_timing()
            full = sum(_domino_sites(t) for t in tilings) * tm["kernel_launches"]
            assert 0 < tm["kernel_sites"] < 0.9 * full, (tm["kernel_sites"], full)
            for k, t in enumerate(tilings):
                want = o.domino_walk(t, steps, seeds[k], P, *rule)
                assert np.array_equal(outs[k], want), (shape, k)
    finally:
        gpu_ctx.set_tile_shape(0)


def test_lozenge_cftp_pairs_on_a_slanted_hexagon_match_oracle(gpu_ctx):
    """AlmostHexDomain(120, 40) (sides of unequal length: the support's
    corners sit off the grid's symmetry), 3 samples, 3 doubling rounds from
    256 on k_pair_lozenge, two samples' chains against the oracle."""
    d = rt.LozengeTiler.AlmostHexDomain(120, 40)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 3, 120)
    gpu_ctx.set_tinymt_params(P)
    masters = [2, 7, 11]
    try:
        gpu_ctx.set_tile_shape(7)
        _say("slanted hexagon lozenge CFTP pairs, 3 samples, 3 rounds")
        tops, bots, _ = gpu_ctx.lozenge_cftp_batch_rounds(tmax, tmin, masters,
                                                          rt.SCHED_DOUBLING, 256, 3)
        tm = gpu_ctx.last_timing()
    finally:
        gpu_ctx.set_tile_shape(0)
    assert tm["shape"] == 7
    assert 0 < tm["kernel_sites"] < 0.97 * len(masters) * N * (N // 3) * tm["kernel_launches"]
    for k in (0, 2):
        seeds = [int(x) for x in o.mt19937(masters[k], 3)][::-1]
        steps = [512, 256, 256]
        assert np.array_equal(tops[k], o.lozenge_walk_segments(tmax, steps, seeds, P)), k
        assert np.array_equal(bots[k], o.lozenge_walk_segments(tmin, steps, seeds, P)), k


def test_single_walks_and_cftp_on_sparse_domains_match_oracle(gpu_ctx):
    """The one-sample entry points (rt_domino_walk, rt_lozenge_walk,
    rt_domino_cftp) leave out the same tiles: an L-shaped domino domain
    (heat bath, the 0.8 toggle), the slanted hexagon, and an Aztec diamond's
    CFTP sample."""
    D = rt.DominoTiler
    t = D.MinTiling(D.RectangleUnion([(0, 0, 200, 12), (188, 0, 12, 200)]))
    N = rt.side(t)
    P = rt.synthetic_params(N * N // 2, 33)
    gpu_ctx.set_tinymt_params(P)
    for rule in (rt.DOMINO_HEATBATH, rt.DOMINO_TOGGLE08):
        _say("L-shaped domino walk, rule", rule)
        got = gpu_ctx.domino_walk(t, 700, 21, rule)
        assert np.array_equal(got, o.domino_walk(t, 700, 21, P, rule))
    d = rt.LozengeTiler.AlmostHexDomain(90, 30)
    tl = rt.LozengeTiler.MinTiling(d)
    NL = rt.side(tl)
    PL = rt.synthetic_params(NL * NL // 3, 34)
    gpu_ctx.set_tinymt_params(PL)
    _say("slanted hexagon lozenge walk")
    assert np.array_equal(gpu_ctx.lozenge_walk(tl, 600, 8), o.lozenge_walk(tl, 600, 8, PL))
    az = D.AztecDiamond(40)
    tmin, tmax = D.MinTiling(az), D.MaxTiling(az)
    NA = rt.side(tmin)
    PA = rt.synthetic_params(NA * NA // 2, 35)
    gpu_ctx.set_tinymt_params(PA)
    _say("Aztec 40 domino CFTP, one sample")
    got, st = gpu_ctx.domino_cftp(tmax, tmin, 6, rt.SCHED_DOUBLING, 128, 12)
    want, wst = o.domino_cftp(tmax, tmin, 6, PA, rt.SCHED_DOUBLING, 128, 12)
    assert np.array_equal(got, want) and st["rounds"] == wst["rounds"]


def test_all_zero_tilings_stay_zero(gpu_ctx):
    """A tiling with no nonzero cell leaves every tile out, so a job of such
    units launches no sweep kernel: walks, batches of only such units and
    with one among others, against the oracle (which runs every cell)."""
    N = 40
    z = np.zeros(N * N, dtype=np.int32)
    P = rt.synthetic_params(N * N // 2, 9)
    gpu_ctx.set_tinymt_params(P)
    _say("all-zero domino tilings")
    assert np.array_equal(o.domino_walk(z, 300, 4, P, rt.DOMINO_HEATBATH), z)
    assert np.array_equal(gpu_ctx.domino_walk(z, 300, 4, rt.DOMINO_HEATBATH), z)
    assert np.array_equal(gpu_ctx.domino_walk(z, 300, 4, rt.DOMINO_TOGGLE08), z)
    for a, b in zip(gpu_ctx.domino_walk_batch([z, z], 300, [4, 5], rt.DOMINO_HEATBATH), [z, z]):
        assert np.array_equal(a, b)
    t = rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(17))
    assert rt.side(t) == N
    outs = gpu_ctx.domino_walk_batch([z, t], 300, [4, 5], rt.DOMINO_HEATBATH)
    assert np.array_equal(outs[0], z)
    assert np.array_equal(outs[1], o.domino_walk(t, 300, 5, P, rt.DOMINO_HEATBATH))
    NL = 27
    zl = np.zeros(NL * NL, dtype=np.int32)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(NL * NL, 10))
    _say("all-zero lozenge tiling")
    assert np.array_equal(gpu_ctx.lozenge_walk(zl, 300, 4), zl)

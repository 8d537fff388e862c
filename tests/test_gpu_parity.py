"""Device path vs the CPU oracle, bit-exact, through the C ABI.

Every case prints what it runs before it launches anything, so a failing
device run names its case.
"""
import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu

RULES = [rt.DOMINO_TOGGLE08, rt.DOMINO_HEATBATH, rt.DOMINO_METRO2P, rt.DOMINO_HEATBATH2P]


def _say(*a):
    print("[case]", *a, flush=True)


def _domino(n):
    d = rt.DominoTiler.AztecDiamond(n)
    return d, rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)


def _params(count, seed=0x5eed):
    return rt.synthetic_params(count, seed)


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("n,steps,seed", [(2, 1, 1), (5, 2, 7), (5, 257, 34893), (16, 1000, 3)])
def test_domino_walk_matches_oracle(gpu_ctx, rule, n, steps, seed):
    d, tmin, tmax = _domino(n)
    N = rt.side(tmin)
    P = _params(N * N // 2)
    gpu_ctx.set_tinymt_params(P)
    _say("domino walk", rule, n, steps, seed)
    got = gpu_ctx.domino_walk(tmin, steps, seed, rule, 0.7, 1.0)
    want = o.domino_walk(tmin, steps, seed, P, rule, 0.7, 1.0)
    assert np.array_equal(got, want)


def test_domino_walk_zero_steps_identity(gpu_ctx):
    d, tmin, _ = _domino(4)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(_params(N * N // 2))
    _say("domino walk 0 steps")
    assert np.array_equal(gpu_ctx.domino_walk(tmin, 0, 5), tmin)


@pytest.mark.parametrize("rule", [rt.DOMINO_TOGGLE08, rt.DOMINO_HEATBATH2P])
def test_domino_walk_segments_matches_oracle(gpu_ctx, rule):
    d, tmin, tmax = _domino(7)
    N = rt.side(tmax)
    P = _params(N * N // 2, seed=11)
    gpu_ctx.set_tinymt_params(P)
    steps, seeds = [256, 128, 33], [4000000001, 17, 2**31 + 5]
    _say("domino segments", rule)
    got = gpu_ctx.domino_walk_segments(tmax, steps, seeds, rule, 0.5, 1.0)
    want = o.domino_walk_segments(tmax, steps, seeds, P, rule, 0.5, 1.0)
    assert np.array_equal(got, want)


def test_domino_weighted_config_rule(gpu_ctx):
    """Two-periodic a=1.0, b=0.5 (the weighted Aztec configuration's
    weights), Metropolis rule, on an order-24 diamond."""
    d, tmin, _ = _domino(24)
    N = rt.side(tmin)
    P = _params(N * N // 2, seed=2)
    gpu_ctx.set_tinymt_params(P)
    _say("domino 2-periodic a=1 b=.5")
    got = gpu_ctx.domino_walk(tmin, 3000, 9, rt.DOMINO_METRO2P, 1.0, 0.5)
    want = o.domino_walk(tmin, 3000, 9, P, rt.DOMINO_METRO2P, 1.0, 0.5)
    assert np.array_equal(got, want)


def test_domino_aztec100_1e4_steps(gpu_ctx):
    """Minimal slice: Walk(MinTiling(AztecDiamond(100)), 10^4, seed=1) with
    both the 0.8 toggle and the 1/2 heat-bath rule, byte for byte."""
    d, tmin, _ = _domino(100)
    N = rt.side(tmin)
    P = _params(N * N // 2)
    gpu_ctx.set_tinymt_params(P)
    o.set_threads(8)
    for rule in (rt.DOMINO_HEATBATH, rt.DOMINO_TOGGLE08):
        _say("aztec-100 1e4", rule)
        got = gpu_ctx.domino_walk(tmin, 10000, 1, rule)
        want = o.domino_walk(tmin, 10000, 1, P, rule)
        assert np.array_equal(got, want)
    o.set_threads(1)


@pytest.mark.parametrize("sched", [rt.SCHED_REFERENCE, rt.SCHED_DOUBLING])
def test_domino_cftp_matches_oracle(gpu_ctx, sched):
    d, tmin, tmax = _domino(5)
    N = rt.side(tmin)
    P = _params(N * N // 2)
    gpu_ctx.set_tinymt_params(P)
    for master in (345903, 1, 2):
        _say("domino cftp", sched, master)
        got, gs = gpu_ctx.domino_cftp(tmax, tmin, master, sched, 128, 40)
        want, ws = o.domino_cftp(tmax, tmin, master, P, sched, 128, 40)
        assert np.array_equal(got, want)
        assert gs["rounds"] == ws["rounds"] and gs["total_steps"] == ws["total_steps"]


def _lozenge(n, m=0):
    d = rt.LozengeTiler.AlmostHexDomain(n, m)
    return d, rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)


@pytest.mark.parametrize("n,m,steps,seed", [(2, 0, 1, 1), (4, 0, 100, 5), (10, 2, 1000, 77),
                                            (20, 0, 2000, 12345)])
def test_lozenge_walk_matches_oracle(gpu_ctx, n, m, steps, seed):
    d, tmin, tmax = _lozenge(n, m)
    N = rt.side(tmin)
    P = _params(N * N)
    gpu_ctx.set_tinymt_params(P)
    _say("lozenge walk", n, m, steps, seed)
    for start in (tmin, tmax):
        got = gpu_ctx.lozenge_walk(start, steps, seed)
        want = o.lozenge_walk(start, steps, seed, P)
        assert np.array_equal(got, want)


def test_lozenge_segments_and_cftp(gpu_ctx):
    d, tmin, tmax = _lozenge(5)
    N = rt.side(tmin)
    P = _params(N * N, seed=3)
    gpu_ctx.set_tinymt_params(P)
    _say("lozenge segments")
    got = gpu_ctx.lozenge_walk_segments(tmin, [64, 10, 300], [9, 8, 7])
    want = o.lozenge_walk_segments(tmin, [64, 10, 300], [9, 8, 7], P)
    assert np.array_equal(got, want)
    for sched in (rt.SCHED_REFERENCE, rt.SCHED_DOUBLING):
        for master in (1, 2, 3):
            _say("lozenge cftp", sched, master)
            g, gs = gpu_ctx.lozenge_cftp(tmax, tmin, master, sched, 64, 40)
            w, ws = o.lozenge_cftp(tmax, tmin, master, P, sched, 64, 40)
            assert np.array_equal(g, w)
            assert gs["rounds"] == ws["rounds"]


@pytest.mark.parametrize("n,steps,seed", [(2, 1, 1), (3, 50, 4), (8, 700, 21)])
def test_recttri_walk_matches_oracle(gpu_ctx, n, steps, seed):
    for start in (rt.RectTriangleTiler.minHex(n), rt.RectTriangleTiler.maxHex(n)):
        N = rt.side(start)
        P = _params(N * N, seed=5)
        gpu_ctx.set_tinymt_params(P)
        _say("recttri walk", n, steps, seed)
        got = gpu_ctx.recttri_walk(start, steps, seed)
        want = o.recttri_walk(start, steps, seed, P)
        assert np.array_equal(got, want)


def test_recttri_weighted_matches_oracle(gpu_ctx):
    start = rt.RectTriangleTiler.minHex(6)
    N = rt.side(start)
    P = _params(N * N, seed=6)
    gpu_ctx.set_tinymt_params(P)
    w = np.array([0.0, 1.0, 0.7, 1.3, 0.5, 2.0, 0.9, 0.5, 2.0, 0.9, 1.1, 0.3, 1.7, 1.1, 0.3, 1.7],
                 dtype=np.float32)
    _say("recttri weighted")
    got = gpu_ctx.recttri_walk(start, 900, 3, w)
    want = o.recttri_walk(start, 900, 3, P, w)
    assert np.array_equal(got, want)


def test_params_too_small_is_rejected(gpu_ctx):
    d, tmin, _ = _domino(5)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(_params(N * N // 2 - 1))
    _say("short param table")
    with pytest.raises(rt.RTError) as e:
        gpu_ctx.domino_walk(tmin, 10, 1)
    assert e.value.code == -2


@pytest.mark.parametrize("engine", [rt.ENGINE_STEP, rt.ENGINE_FUSED, rt.ENGINE_FUSED_SWAR])
def test_engines_agree_with_oracle(gpu_ctx, engine):
    """Both kernel shapes (one launch per step; fused LDS sweep blocks) on
    domains wider and taller than one fused tile."""
    gpu_ctx.set_engine(engine)
    try:
        d, tmin, tmax = _domino(60)
        N = rt.side(tmin)
        P = _params(N * N // 2, seed=8)
        gpu_ctx.set_tinymt_params(P)
        for rule in RULES:
            _say("engine", engine, "domino-60 rule", rule)
            got = gpu_ctx.domino_walk(tmax, 777, 4, rule, 0.7, 1.0)
            want = o.domino_walk(tmax, 777, 4, P, rule, 0.7, 1.0)
            assert np.array_equal(got, want)
        dl, lmin, lmax = _lozenge(40, 3)
        NL = rt.side(lmin)
        PL = _params(NL * NL, seed=9)
        gpu_ctx.set_tinymt_params(PL)
        _say("engine", engine, "lozenge-40")
        assert np.array_equal(gpu_ctx.lozenge_walk(lmin, 999, 6), o.lozenge_walk(lmin, 999, 6, PL))
        r = rt.RectTriangleTiler.maxHex(30)
        NR = rt.side(r)
        PR = _params(NR * NR, seed=10)
        gpu_ctx.set_tinymt_params(PR)
        _say("engine", engine, "recttri-30")
        assert np.array_equal(gpu_ctx.recttri_walk(r, 555, 2), o.recttri_walk(r, 555, 2, PR))
    finally:
        gpu_ctx.set_engine(rt.ENGINE_AUTO)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7])
def test_tile_shapes_agree_with_oracle(gpu_ctx, shape):
    """Every sweep-block tile shape (64 rows / 8 waves, 64 rows / 16 waves,
    32 rows / 16 waves; pairs split; 64 rows / 16 waves at eight waves per
    SIMD) on domains spanning several tiles, for a walk of each byte-cell
    model and a batched CFTP of each."""
    gpu_ctx.set_tile_shape(shape)
    try:
        d, tmin, tmax = _domino(45)
        N = rt.side(tmin)
        P = _params(N * N // 2, seed=12)
        gpu_ctx.set_tinymt_params(P)
        _say("shape", shape, "domino-45")
        assert np.array_equal(gpu_ctx.domino_walk(tmin, 901, 5, rt.DOMINO_HEATBATH),
                              o.domino_walk(tmin, 901, 5, P, rt.DOMINO_HEATBATH))
        dl, lmin, lmax = _lozenge(33, 4)
        NL = rt.side(lmin)
        PL = _params(NL * NL, seed=13)
        gpu_ctx.set_tinymt_params(PL)
        _say("shape", shape, "lozenge-33")
        assert np.array_equal(gpu_ctx.lozenge_walk(lmin, 1201, 8), o.lozenge_walk(lmin, 1201, 8, PL))
        dl, lmin, lmax = _lozenge(6)
        NL = rt.side(lmin)
        PL = _params(NL * NL, seed=14)
        gpu_ctx.set_tinymt_params(PL)
        _say("shape", shape, "lozenge-6 cftp batch")
        outs, stats = gpu_ctx.lozenge_cftp_batch(lmax, lmin, [4, 5, 6], rt.SCHED_DOUBLING, 64, 40)
        for k, master in enumerate([4, 5, 6]):
            w, ws = o.lozenge_cftp(lmax, lmin, master, PL, rt.SCHED_DOUBLING, 64, 40)
            assert np.array_equal(outs[k], w)
            assert stats[k]["rounds"] == ws["rounds"]
        d, tmin, tmax = _domino(7)
        N = rt.side(tmin)
        P = _params(N * N // 2, seed=15)
        gpu_ctx.set_tinymt_params(P)
        _say("shape", shape, "domino-7 cftp batch")
        outs, stats = gpu_ctx.domino_cftp_batch(tmax, tmin, [2, 3], rt.SCHED_DOUBLING, 64, 30)
        for k, master in enumerate([2, 3]):
            w, ws = o.domino_cftp(tmax, tmin, master, P, rt.SCHED_DOUBLING, 64, 30)
            assert np.array_equal(outs[k], w)
            assert stats[k]["rounds"] == ws["rounds"]
    finally:
        gpu_ctx.set_tile_shape(0)

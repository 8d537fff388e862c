"""The allocation-free domino walk (rt_domino_walk_noalloc) against the
ordinary entry point (rt_domino_walk) through the C ABI: bit for bit, on two
calls in a row, the second on a smaller input (the fixed device array keeps
the first call's contents), for both flip rules of the plain Walk; and the
size limit."""
import numpy as np
import pytest

import randomtilingspublic_amd as rt

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("rule", [rt.DOMINO_TOGGLE08, rt.DOMINO_HEATBATH, rt.DOMINO_METRO2P])
def test_noalloc_walk_equals_walk_on_two_calls(gpu_ctx, rule):
    big = rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(100))
    small = rt.DominoTiler.MaxTiling(rt.DominoTiler.AztecDiamond(13))
    N = rt.side(big)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, 41))
    gpu_ctx.set_engine(rt.ENGINE_AUTO)
    for t, steps, seed in ((big, 3000, 1), (small, 777, 12345)):
        _say("noalloc domino walk", rule, rt.side(t), steps, seed)
        got = gpu_ctx.domino_walk_noalloc(t, steps, seed, rule, 1.0, 0.5)
        want = gpu_ctx.domino_walk(t, steps, seed, rule, 1.0, 0.5)
        assert np.array_equal(got, want)
        assert not np.array_equal(got, t)
    _say("noalloc zero steps")
    assert np.array_equal(gpu_ctx.domino_walk_noalloc(small, 0, 3, rule), small)


def test_noalloc_walk_size_limit(gpu_ctx):
    # N = 1136 (Aztec order 565) fits the 16 MiB array, N = 1138 does not;
    # the check comes before any device work.
    t = rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(566))
    N = rt.side(t)
    assert N == 1138
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
    _say("noalloc size limit", N)
    with pytest.raises(rt.RTError) as e:
        gpu_ctx.domino_walk_noalloc(t, 10, 1, rt.DOMINO_HEATBATH)
    assert e.value.code == -1
    t2 = rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(565))
    assert rt.side(t2) == 1136
    _say("noalloc largest side", rt.side(t2))
    got = gpu_ctx.domino_walk_noalloc(t2, 40, 2, rt.DOMINO_HEATBATH)
    assert np.array_equal(got, gpu_ctx.domino_walk(t2, 40, 2, rt.DOMINO_HEATBATH))

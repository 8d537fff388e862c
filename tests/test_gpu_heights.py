"""Height functions computed on the device (rt_domino_walk_batch_heights)
against the reference's TilingToHeightfunc (DominoTiler.cpp:555-593): its
own outputs on the minimal tilings of Aztec diamonds
(tests/golden/reference_host.npz, written by the reference's compiled host
code), and the host restatement (itself checked against the reference
library in test_host.py) on walked tilings of uneven domains."""
import os

import numpy as np
import pytest

import randomtilingspublic_amd as rt

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_host.npz")
D = rt.DominoTiler


def _say(*a):
    print("[case]", *a, flush=True)


def test_device_heights_equal_reference_vectors(gpu_ctx):
    g = np.load(GOLDEN)
    ns = (1, 2, 3, 7, 16, 33)
    tilings = [g[f"domino_aztec{n}_min"] for n in ns]
    for n, t in zip(ns, tilings):
        assert np.array_equal(D.TilingToDomain(t), g[f"domino_aztec{n}_domain"])
    big = max(rt.side(t) for t in tilings)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(big * big // 2, 50))
    _say("device heights of Aztec minimal tilings", ns)
    outs, _, hs = gpu_ctx.domino_walk_batch_heights(tilings, 0, list(range(len(ns))),
                                                    rt.DOMINO_HEATBATH)
    for n, t, o, h in zip(ns, tilings, outs, hs):
        assert np.array_equal(o, t)
        assert np.array_equal(h, g[f"domino_aztec{n}_min_hf"]), n


def _domains():
    ts = [D.MinTiling(D.AztecDiamond(n)) for n in (4, 30, 61)]
    # a U: two arms joined only through the base, so the search must go down
    # one arm and up the other
    ts.append(D.MinTiling(D.RectangleUnion([(0, 0, 40, 10), (0, 30, 40, 10), (30, 0, 10, 40)])))
    ts.append(D.MaxTiling(D.RectangleUnion([(0, 0, 20, 30), (10, 24, 30, 12), (36, 4, 6, 40)])))
    return ts


@pytest.mark.parametrize("use_graph", [False, True])
def test_device_heights_of_walked_tilings(gpu_ctx, use_graph):
    ts = _domains()
    big = max(rt.side(t) for t in ts)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(big * big // 2, 51))
    seeds = [11, 12, 13, 14, 15]
    _say("device heights after 900 heat-bath steps, graph", use_graph)
    outs, counts, hs = gpu_ctx.domino_walk_batch_heights(ts, 900, seeds, rt.DOMINO_HEATBATH,
                                                         use_graph=use_graph)
    plain = gpu_ctx.domino_walk_batch(ts, 900, seeds, rt.DOMINO_HEATBATH)
    for t0, t, p, h, c in zip(ts, outs, plain, hs, counts):
        assert np.array_equal(t, p)
        assert not np.array_equal(t, t0)
        want = D.TilingToHeightfunc(t, D.TilingToDomain(t))
        assert np.array_equal(h, want)
        assert list(c) == [int(((t >> b) & 1).sum()) for b in range(4)] + [
            int(((t == 3) | (t == 12)).sum())]
    _say("device heights without downloading the tilings")
    none, _, hs2 = gpu_ctx.domino_walk_batch_heights(ts, 900, seeds, rt.DOMINO_HEATBATH,
                                                     use_graph=use_graph, download=False)
    assert none is None
    for a, b in zip(hs, hs2):
        assert np.array_equal(a, b)


def test_device_heights_large_aztec(gpu_ctx):
    t = D.MinTiling(D.AztecDiamond(300))
    N = rt.side(t)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, 52))
    _say("device heights Aztec 300 after 3000 steps")
    outs, _, hs = gpu_ctx.domino_walk_batch_heights([t, t], 3000, [1, 2], rt.DOMINO_TOGGLE08)
    for o, h in zip(outs, hs):
        assert np.array_equal(h, D.TilingToHeightfunc(o, D.AztecDiamond(300)))

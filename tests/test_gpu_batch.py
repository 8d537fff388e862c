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

"""Edge cases of the batched paths on the device: a CFTP batch that cannot
coalesce within its round cap fails with ENOCOUPLE and leaves the context
usable; zero-step and one-chain batch walks; malformed batch inputs; the
XCD tile order, the halo depth; the CFTP round callback; forced tile
shapes under the four-cells-per-word lozenge kernel."""
import sys

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def test_cftp_batch_round_cap_then_context_still_works(gpu_ctx):
    d = rt.LozengeTiler.AlmostHexDomain(20, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 31)
    gpu_ctx.set_tinymt_params(P)
    _say("lozenge cftp batch capped at 2 rounds")
    with pytest.raises(rt.RTError) as e:
        gpu_ctx.lozenge_cftp_batch(tmax, tmin, [1, 2], rt.SCHED_DOUBLING, 4, 2)
    assert e.value.code == -4  # ENOCOUPLE
    _say("lozenge cftp batch after the failure")
    outs, stats = gpu_ctx.lozenge_cftp_batch(tmax, tmin, [1, 2], rt.SCHED_DOUBLING, 1024, 20)
    for k, m in enumerate((1, 2)):
        want, _ = o.lozenge_cftp(tmax, tmin, m, P, rt.SCHED_DOUBLING, 1024, 20)
        assert np.array_equal(outs[k], want)


def test_walk_batch_zero_steps_and_single_chain(gpu_ctx):
    t = rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(20))
    N = rt.side(t)
    P = rt.synthetic_params(N * N // 2, 32)
    gpu_ctx.set_tinymt_params(P)
    _say("walk batch zero steps")
    outs = gpu_ctx.domino_walk_batch([t, t], 0, [1, 2], rt.DOMINO_HEATBATH)
    assert all(np.array_equal(x, t) for x in outs)
    _say("walk batch one chain")
    (one,) = gpu_ctx.domino_walk_batch([t], 333, [9], rt.DOMINO_HEATBATH, use_graph=True)
    assert np.array_equal(one, o.domino_walk(t, 333, 9, P, rt.DOMINO_HEATBATH))


def test_walk_batch_rejects_bad_chains(gpu_ctx):
    gpu_ctx.set_tinymt_params(rt.synthetic_params(4096, 33))
    _say("walk batch odd side")
    with pytest.raises(rt.RTError):
        gpu_ctx.domino_walk_batch([np.zeros(49, np.int32)], 10, [1])
    with pytest.raises(rt.RTError):
        gpu_ctx.domino_walk_batch([rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(60))],
                                  10, [1])  # 126 x 63 sites > 4096 parameter sets


def test_xcd_order_never_changes_results(gpu_ctx):
    ts = [rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(n)) for n in (40, 70, 9)]
    big = max(rt.side(t) for t in ts)
    P = rt.synthetic_params(big * big // 2, 34)
    gpu_ctx.set_tinymt_params(P)
    outs = []
    try:
        for on in (0, 1):
            _say("walk batch xcd order", on)
            gpu_ctx.set_xcd_order(on)
            outs.append(gpu_ctx.domino_walk_batch(ts, 500, [3, 4, 5], rt.DOMINO_HEATBATH))
    finally:
        gpu_ctx.set_xcd_order(1)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert np.array_equal(outs[1][2], o.domino_walk(ts[2], 500, 5, P, rt.DOMINO_HEATBATH))


def test_xcd_order_never_changes_pair_or_bibone_results(gpu_ctx):
    """Mode 2 (XCD order on every launch) against mode 0 on the lozenge pair
    kernel (tile shape 7, more workgroups than CUs, the remap of large
    benchmark grids) and on the bibone edge kernel (a tile shape with more
    tiles than CUs)."""
    d = rt.LozengeTiler.AlmostHexDomain(60, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N, 35))
    masters = list(range(1, 13))
    outs = []
    try:
        gpu_ctx.set_tile_shape(7)
        for on in (0, 2):
            _say("pair cftp rounds xcd order", on)
            gpu_ctx.set_xcd_order(on)
            tops, bots, _ = gpu_ctx.lozenge_cftp_batch_rounds(tmax, tmin, masters,
                                                              rt.SCHED_DOUBLING, 256, 3)
            outs.append((tops, bots))
    finally:
        gpu_ctx.set_tile_shape(0)
        gpu_ctx.set_xcd_order(1)
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    T = rt.TriangleDimerTiler
    t = T.Hexagon(70, 71, 70)
    Nb = rt.side(t)
    gpu_ctx.set_tinymt_params(rt.synthetic_params((Nb // 2) ** 2, 36))
    got = []
    try:
        gpu_ctx.set_tile_shape(6)  # 48 x 48 tiles, halo 12: 7 x 7 tiles here
        for on in (0, 2):
            _say("bibone walk xcd order", on)
            gpu_ctx.set_xcd_order(on)
            got.append(gpu_ctx.bibone_walk(t, 700, 3))
    finally:
        gpu_ctx.set_tile_shape(0)
        gpu_ctx.set_xcd_order(1)
    assert np.array_equal(got[0], got[1])
    assert not np.array_equal(got[0], np.asarray(t))


def test_round_callback_reports_every_round(gpu_ctx):
    d = rt.LozengeTiler.AlmostHexDomain(12, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N, 35))
    calls = []
    gpu_ctx.set_round_callback(lambda r, n, fin, tot: calls.append((r, n, fin, tot)))
    try:
        _say("lozenge cftp batch with a round callback")
        outs, stats = gpu_ctx.lozenge_cftp_batch(tmax, tmin, [1, 2, 3], rt.SCHED_DOUBLING, 64, 20)
        R = max(s["rounds"] for s in stats)
        assert [c[0] for c in calls] == list(range(1, R + 1))
        assert [c[1] for c in calls] == [64 << (r - 1) for r in range(1, R + 1)]
        assert all(c[3] == 3 for c in calls)
        for r, _, fin, _ in calls:
            assert fin == sum(1 for s in stats if s["rounds"] <= r)
        calls.clear()
        _say("lozenge cftp one sample with a round callback")
        out, st = gpu_ctx.lozenge_cftp(tmax, tmin, 2, rt.SCHED_DOUBLING, 64, 20)
        assert np.array_equal(out, outs[1])
        assert [c[0] for c in calls] == list(range(1, st["rounds"] + 1))
        assert [c[2] for c in calls] == [0] * (st["rounds"] - 1) + [1]
        calls.clear()
        gpu_ctx.set_round_callback(None)
        gpu_ctx.lozenge_cftp(tmax, tmin, 2, rt.SCHED_DOUBLING, 64, 20)
        assert calls == []
    finally:
        gpu_ctx.set_round_callback(None)


@pytest.mark.parametrize("halo", [2, 6, 9, 12])
def test_halo_depth_never_changes_results(gpu_ctx, halo):
    # Domains wide enough for two tile columns at either halo (64-column
    # tiles overlap by 2 * halo): the hexagon of side 100 (69 lozenge
    # columns) and the Aztec diamond of order 70 (73 domino columns).
    d30 = rt.LozengeTiler.AlmostHexDomain(30, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d30), rt.LozengeTiler.MaxTiling(d30)
    big = rt.LozengeTiler.MinTiling(rt.LozengeTiler.AlmostHexDomain(100, 0))
    P = rt.synthetic_params(rt.side(big) ** 2, 36)
    want_c = o.lozenge_cftp(tmax, tmin, 4, P, rt.SCHED_DOUBLING, 256, 20)[0]
    want_w = o.lozenge_walk(big, 300, 8, P)
    ts = [rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(n)) for n in (70, 12)]
    Pd = rt.synthetic_params(rt.side(ts[0]) ** 2 // 2, 37)
    want_d = [o.domino_walk(t, 700, s, Pd, rt.DOMINO_HEATBATH) for t, s in zip(ts, (6, 7))]
    gpu_ctx.set_tinymt_params(P)
    try:
        gpu_ctx.set_halo(halo)
        for shape in (2, 6, 7):
            _say("lozenge cftp batch halo", halo, "shape", shape)
            gpu_ctx.set_tile_shape(shape)
            outs, _ = gpu_ctx.lozenge_cftp_batch(tmax, tmin, [4], rt.SCHED_DOUBLING, 256, 20)
            assert np.array_equal(outs[0], want_c)
        gpu_ctx.set_tile_shape(0)
        _say("lozenge walk halo", halo)
        assert np.array_equal(gpu_ctx.lozenge_walk(big, 300, 8), want_w)
        gpu_ctx.set_tinymt_params(Pd)
        for shape in (1, 7):
            gpu_ctx.set_tile_shape(shape)
            _say("domino walk batch halo", halo, "shape", shape)
            outs = gpu_ctx.domino_walk_batch(ts, 700, [6, 7], rt.DOMINO_HEATBATH)
            for got, w in zip(outs, want_d):
                assert np.array_equal(got, w)
    finally:
        gpu_ctx.set_halo(0)
        gpu_ctx.set_tile_shape(0)
    with pytest.raises(rt.RTError):
        gpu_ctx.set_halo(-1)
    with pytest.raises(rt.RTError):
        gpu_ctx.set_halo(13)


@pytest.mark.parametrize("halo", [3, 9])
def test_swar_engine_halo_depth_never_changes_results(gpu_ctx, halo):
    # The four-cells-per-word lozenge kernel at halo depths other than 6,
    # on the side-100 hexagon (two tile columns at either depth).
    d = rt.LozengeTiler.AlmostHexDomain(100, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 39)
    want = o.lozenge_walk(tmin, 300, 8, P)
    steps, seeds = [96, 64], [11, 12]
    want_seg = o.lozenge_walk_segments(tmax, steps, seeds, P)
    gpu_ctx.set_tinymt_params(P)
    try:
        gpu_ctx.set_engine(rt.ENGINE_FUSED_SWAR)
        gpu_ctx.set_halo(halo)
        for shape in (2, 3):
            gpu_ctx.set_tile_shape(shape)
            _say("swar lozenge walk halo", halo, "shape", shape)
            assert np.array_equal(gpu_ctx.lozenge_walk(tmin, 300, 8), want)
            assert np.array_equal(gpu_ctx.lozenge_walk_segments(tmax, steps, seeds), want_seg)
    finally:
        gpu_ctx.set_halo(0)
        gpu_ctx.set_engine(rt.ENGINE_AUTO)
        gpu_ctx.set_tile_shape(0)


def test_domino_cftp_rejects_non_monotone_rules(gpu_ctx):
    # The 0.8 toggle and the two-periodic Metropolis rule are not monotone
    # couplings: coalescence of the extreme chains would not be exact.
    d = rt.DominoTiler.AztecDiamond(4)
    tmin, tmax = rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, 40))
    _say("domino cftp rule checks")
    for rule in (rt.DOMINO_TOGGLE08, rt.DOMINO_METRO2P, 7, -1):
        with pytest.raises(rt.RTError) as e:
            gpu_ctx.domino_cftp(tmax, tmin, 1, rt.SCHED_DOUBLING, 64, 10, rule)
        assert e.value.code == -1
        with pytest.raises(rt.RTError) as e:
            gpu_ctx.domino_cftp_batch(tmax, tmin, [1, 2], rt.SCHED_DOUBLING, 64, 10, rule)
        assert e.value.code == -1
    srv = rt.Server(gpu_ctx, 64 << 20, max_active=4)
    try:
        for rule in (rt.DOMINO_TOGGLE08, rt.DOMINO_METRO2P, 9):
            with pytest.raises(rt.RTError) as e:
                srv.submit(rt.MODEL_DOMINO, tmax, tmin, 1, rt.SCHED_DOUBLING, 64, 10, rule)
            assert e.value.code == -1
        rid = srv.submit(rt.MODEL_DOMINO, tmax, tmin, 1, rt.SCHED_DOUBLING, 64, 10,
                         rt.DOMINO_HEATBATH2P, 1.0, 0.5)
        srv.wait(rid, 60.0)
        assert srv.poll(rid)[0] == 0
    finally:
        srv.close()
    for rule in (rt.DOMINO_HEATBATH, rt.DOMINO_HEATBATH2P):
        gpu_ctx.domino_cftp(tmax, tmin, 1, rt.SCHED_DOUBLING, 64, 10, rule)


@pytest.mark.parametrize("shape", [0, 2, 3, 6, 7])
def test_swar_engine_cftp_batch_with_forced_shapes(gpu_ctx, shape):
    # The four-cells-per-word lozenge kernel has 64- and 32-row tiles only;
    # a forced shape without a SWAR kernel (7) falls back to 64-row tiles.
    d = rt.LozengeTiler.AlmostHexDomain(12, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, 38)
    gpu_ctx.set_tinymt_params(P)
    try:
        gpu_ctx.set_engine(rt.ENGINE_FUSED_SWAR)
        gpu_ctx.set_tile_shape(shape)
        _say("swar lozenge cftp batch shape", shape)
        outs, _ = gpu_ctx.lozenge_cftp_batch(tmax, tmin, [2, 3], rt.SCHED_DOUBLING, 256, 20)
        for k, m in enumerate((2, 3)):
            want, _ = o.lozenge_cftp(tmax, tmin, m, P, rt.SCHED_DOUBLING, 256, 20)
            assert np.array_equal(outs[k], want)
        if shape == 7:
            assert gpu_ctx.last_timing()["shape"] == 2
    finally:
        gpu_ctx.set_engine(rt.ENGINE_AUTO)
        gpu_ctx.set_tile_shape(0)

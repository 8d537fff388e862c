"""Bibone walk on the GPU against the oracle: bit-exact lozenge tilings for
the same TinyMT parameters, seed and step count (TriangleDimerTiler::Walk,
TriangleDimerTiler.cpp:101-695), from the reference's IceCreamCone starts
and from hexagons, on all three engines (edge form, eight-plane sweep
blocks, one launch per reference kernel), every edge-form tile shape, and a
start that the edge form does not take; plus the device result's own
invariants at a larger size."""
import sys

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu
T = rt.TriangleDimerTiler


def _say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("engine", [rt.ENGINE_AUTO, rt.ENGINE_FUSED, rt.ENGINE_STEP])
@pytest.mark.parametrize("n,m,steps,seed", [(6, 0, 0, 1), (6, 0, 1, 2), (6, 0, 64, 3),
                                            (9, 2, 500, 11), (13, 5, 1500, 4), (30, 0, 3000, 77),
                                            (70, 9, 1200, 123), (150, 20, 700, 5)])
def test_bibone_walk_matches_oracle(gpu_ctx, n, m, steps, seed, engine):
    """ENGINE_AUTO: the edge form (one bit per lattice edge, a cycle flip
    per move, 8 steps per launch on 64 x 64 tiles; the cone of side 150 has
    5 x 5 tiles); ENGINE_FUSED: the eight planes in sweep blocks; ENGINE_STEP:
    one launch per reference kernel."""
    t = T.IceCreamCone(n, m)
    N = rt.side(t)
    P = rt.synthetic_params((N // 2) ** 2, seed=seed + 1)
    gpu_ctx.set_tinymt_params(P)
    _say("bibone walk", n, m, steps, seed, "engine", engine)
    try:
        gpu_ctx.set_engine(engine)
        got = gpu_ctx.bibone_walk(t, steps, seed)
    finally:
        gpu_ctx.set_engine(rt.ENGINE_AUTO)
    want = o.bibone_walk(t, steps, seed, P)
    assert np.array_equal(got, want)
    tm = gpu_ctx.last_timing()
    assert tm["steps"] == steps and tm["attempts"] == steps * (N // 2) ** 2
    assert (tm["shape"] != 0) == (engine == rt.ENGINE_AUTO)


def test_bibone_walk_large_cone_invariants(gpu_ctx):
    t = T.IceCreamCone(200, 0)
    N = rt.side(t)
    gpu_ctx.set_tinymt_params(rt.synthetic_params((N // 2) ** 2, seed=9))
    _say("bibone walk cone 200")
    got = gpu_ctx.bibone_walk(t, 4000, 5)
    assert np.array_equal(T.TilingToDomain(got), T.TilingToDomain(t))
    assert (got != t).sum() > 100
    # walking on from the result in two halves equals one walk only if the
    # second half restarts its schedule: check determinism instead
    again = gpu_ctx.bibone_walk(t, 4000, 5)
    assert np.array_equal(got, again)


def test_bibone_walk_rejects_bad_input(gpu_ctx):
    gpu_ctx.set_tinymt_params(rt.synthetic_params(16))
    _say("bibone bad input")
    with pytest.raises(rt.RTError):
        gpu_ctx.bibone_walk(np.zeros(49, np.int32), 1, 1)  # odd side
    with pytest.raises(rt.RTError) as e:
        gpu_ctx.bibone_walk(T.IceCreamCone(6, 0), 1, 1)  # 16 < 10*10 parameter sets
    assert e.value.code == -2


@pytest.mark.parametrize("a,b,c,steps,seed", [(5, 5, 6, 700, 3), (20, 20, 21, 1500, 8)])
def test_bibone_walk_from_hexagon_matches_oracle(gpu_ctx, a, b, c, steps, seed):
    t = T.Hexagon(a, b, c)
    N = rt.side(t)
    P = rt.synthetic_params((N // 2) ** 2, seed=seed + 5)
    gpu_ctx.set_tinymt_params(P)
    _say("bibone walk hexagon", a, b, c, steps, seed)
    got = gpu_ctx.bibone_walk(t, steps, seed)
    want = o.bibone_walk(t, steps, seed, P)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7])
def test_bibone_edge_tile_shapes_match_oracle(gpu_ctx, shape):
    """Every edge-form tile geometry (32/40/48/64-cell tiles, halo 4-16, so
    2-8 steps per launch) on the hexagon 20, 19, 20 (44 x 44 plane cells)."""
    t = T.Hexagon(20, 19, 20)
    N = rt.side(t)
    P = rt.synthetic_params((N // 2) ** 2, seed=31)
    gpu_ctx.set_tinymt_params(P)
    _say("bibone edge tile shape", shape)
    try:
        gpu_ctx.set_tile_shape(shape)
        got = gpu_ctx.bibone_walk(t, 1001, 6)
        assert gpu_ctx.last_timing()["shape"] == shape
    finally:
        gpu_ctx.set_tile_shape(0)
    assert np.array_equal(got, o.bibone_walk(t, 1001, 6, P))


def test_bibone_walk_start_off_the_edge_form(gpu_ctx):
    """A hexagon moved up by one vertex row touches the outer ring of cells
    that the reference never refreshes: the edge form refuses it and the
    walk runs on the eight planes, still equal to the oracle."""
    t = np.asarray(T.Hexagon(6, 6, 7))
    N = rt.side(t)
    t = np.roll(t.reshape(N, N), -2, axis=0).ravel()
    assert T.EdgeForm(t) is None
    P = rt.synthetic_params((N // 2) ** 2, seed=32)
    gpu_ctx.set_tinymt_params(P)
    _say("bibone start off the edge form")
    got = gpu_ctx.bibone_walk(t, 900, 4)
    assert gpu_ctx.last_timing()["shape"] == 0
    assert np.array_equal(got, o.bibone_walk(t, 900, 4, P))

"""Rectangle-triangle walk (RectTriangleTiler::Walk, recttrianglekernel.cl:
37-117) with unit weights: the oracle's chain visits the same set of
tilings from every seed and is uniform on it (its moves are symmetric:
2-cycles for rectangle pairs, a symmetric 4-way choice for the five
all-triangle / single-rectangle hexagon states), checked by chi-square on
the side-1 and side-2 hexagons.  A wrong move-table entry, index or
threshold breaks the symmetry or the reachable set."""
import collections

import numpy as np
from scipy import stats

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

R = rt.RectTriangleTiler


def _law(n, samples, steps, pseed):
    t0 = R.minHex(n)
    N = rt.side(t0)
    P = rt.synthetic_params(N * N, pseed)
    cnt = collections.Counter()
    for s in range(1, samples + 1):
        t = o.recttri_walk(t0, steps, s, P)
        cnt[tuple(np.asarray(t))] += 1
    return t0, cnt


def test_side1_hexagon_walk_is_uniform_on_five_tilings():
    t0, cnt = _law(1, 3000, 40, 5)
    assert len(cnt) == 5
    assert tuple(np.asarray(R.maxHex(1))) in cnt and tuple(np.asarray(t0)) in cnt
    assert stats.chisquare(list(cnt.values())).pvalue > 1e-3


def test_side2_hexagon_walk_is_uniform():
    t0, cnt = _law(2, 4000, 200, 6)
    assert len(cnt) == 368  # every reachable tiling was seen
    assert tuple(np.asarray(R.maxHex(2))) in cnt
    # slopeHex(2) is never reached from minHex(2) by the walk
    assert tuple(np.asarray(R.slopeHex(2))) not in cnt
    assert stats.chisquare(list(cnt.values())).pvalue > 1e-3

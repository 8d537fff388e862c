"""Bibone hexagon constructor (no reference counterpart): its vertex set is
the hexagon, and its tiling decodes to a perfect matching of it.  The
decoder below reads the three dimer-edge families back out of the tiling
cells (the cell layout of TriangleDimerTiler.cpp:738-753) and is itself
pinned on the reference's IceCreamCone starts."""
import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

T = rt.TriangleDimerTiler


def decode(t):
    """(h, l, r) on the E x E vertex grid: h(i,j) = (i,j)-(i,j+1),
    l(i,j) = (i,j)-(i+1,j), r(i,j) = (i,j+1)-(i+1,j), read from the H cells
    (even row, odd col: bit 4 = r(a,b), bit 8 = l(a,b)) and the R cells
    (odd, odd: bit 1 = h(a,b))."""
    N = rt.side(t)
    E = N // 2
    g = np.asarray(t).reshape(N, N)
    h = np.zeros((E, E), int)
    l = np.zeros((E, E), int)
    r = np.zeros((E, E), int)
    for a in range(1, E):
        for b in range(1, E - 1):
            hc = g[2 * a, 2 * b + 1]
            r[a, b] = (hc >> 2) & 1
            l[a, b] = (hc >> 3) & 1
            h[a, b] = g[2 * a + 1, 2 * b + 1] & 1
    return h, l, r


def encode(h, l, r):
    """The cell states of every H, L and R cell from the three families."""
    E = h.shape[0]
    N = 2 * E
    t = np.zeros((N, N), int)
    H = lambda a, b: h[a, b] if 0 <= a < E and 0 <= b < E else 0
    L = lambda a, b: l[a, b] if 0 <= a < E and 0 <= b < E else 0
    R = lambda a, b: r[a, b] if 0 <= a < E and 0 <= b < E else 0
    for i in range(3, N - 1):
        a = i // 2
        for j in range(3, N - 1):
            b = j // 2
            if i % 2 == 0:
                if j % 2:
                    t[i, j] = R(a - 1, b) | L(a - 1, b + 1) << 1 | R(a, b) << 2 | L(a, b) << 3
            elif j % 2 == 0:
                t[i, j] = H(a, b) | R(a, b) << 1 | H(a + 1, b - 1) << 2 | R(a, b - 1) << 3
            else:
                t[i, j] = H(a, b) | L(a, b + 1) << 1 | H(a + 1, b) << 2 | L(a, b) << 3
    return t.ravel()


def degrees(h, l, r):
    d = h.copy()
    d[:, 1:] += h[:, :-1]           # (i, j-1)-(i, j)
    d += l
    d[1:, :] += l[:-1, :]           # (i-1, j)-(i, j)
    d[:, 1:] += r[:, :-1]           # r(i, j-1) = (i, j)-(i+1, j-1)
    d[1:, :] += r[:-1, :]           # r(i-1, j) = (i-1, j+1)-(i, j)
    return d


def hexagon_vertices(a, b, c, E):
    v = np.zeros((E, E), int)
    for i in range(b + c + 1):
        lo, hi = -min(i, b), a - max(0, i - c)
        v[i + 2, lo + 2 + b:hi + 3 + b] = 1
    return v


@pytest.mark.parametrize("n,m", [(6, 0), (9, 2), (13, 5), (30, 0)])
def test_decoder_on_reference_cones(n, m):
    t = T.IceCreamCone(n, m)
    h, l, r = decode(t)
    assert np.array_equal(encode(h, l, r), t)
    v = np.asarray(T.TilingToVertices(t)).reshape(h.shape[0] + 1, -1)[:-1, :-1]
    assert np.array_equal(degrees(h, l, r), v)  # every vertex of the cone covered once


@pytest.mark.parametrize("a,b,c", [(1, 1, 2), (3, 3, 4), (4, 3, 3), (2, 5, 1), (7, 7, 8),
                                   (10, 4, 9), (0, 3, 3)])
def test_hexagon_is_a_perfect_matching_of_the_hexagon(a, b, c):
    t = T.Hexagon(a, b, c)
    N = rt.side(t)
    E = N // 2
    h, l, r = decode(t)
    assert np.array_equal(encode(h, l, r), t)
    want = hexagon_vertices(a, b, c, E)
    assert np.array_equal(degrees(h, l, r), want)
    v = np.asarray(T.TilingToVertices(t)).reshape(E + 1, E + 1)[:-1, :-1]
    assert np.array_equal(v, want)
    assert want.sum() == a * b + b * c + c * a + a + b + c + 1


def test_odd_hexagon_is_rejected():
    with pytest.raises(rt.RTError):
        T.Hexagon(3, 3, 3)  # 37 vertices


def test_hexagon_walk_keeps_the_domain():
    t = T.Hexagon(6, 6, 7)
    N = rt.side(t)
    P = rt.synthetic_params((N // 2) ** 2, seed=4)
    got = o.bibone_walk(t, 600, 9, P)
    assert np.array_equal(T.TilingToDomain(got), T.TilingToDomain(t))
    h, l, r = decode(got)
    assert np.array_equal(degrees(h, l, r), hexagon_vertices(6, 6, 7, N // 2))
    assert (got != t).sum() > 20


@pytest.mark.parametrize("start", ["cone9_2", "cone30_0", "hex7_7_8", "hex10_4_9", "butterfly1"])
def test_edge_form_is_the_decoded_matching(start):
    """rt_bibone_edge_form (the start of the walk's edge-form engine) gives
    the three dimer-edge families that decode() reads off the cells."""
    kind, *args = start.replace("cone", "cone_").replace("hex", "hex_").replace(
        "butterfly", "bf_").split("_")
    args = [int(x) for x in args]
    t = {"cone": T.IceCreamCone, "hex": T.Hexagon, "bf": T.TestButterfly}[kind](*args)
    e = T.EdgeForm(t)
    assert e is not None
    h, l, r = decode(t)
    assert np.array_equal(e, h | l << 1 | r << 2)


def test_edge_form_refuses_starts_it_does_not_reproduce():
    t = np.asarray(T.Hexagon(6, 6, 7))
    N = rt.side(t)
    # a covered vertex on the second vertex row: the reference's unrefreshed
    # outer ring of cells would see it
    up = np.roll(t.reshape(N, N), -2, axis=0).ravel()
    assert T.EdgeForm(up) is None
    # two lozenge cells that disagree on a shared edge
    bad = t.copy().reshape(N, N)
    i, j = np.argwhere(bad[0::2, 1::2] != 0)[0]
    bad[2 * i, 2 * j + 1] ^= 1
    assert T.EdgeForm(bad.ravel()) is None
    assert T.EdgeForm(t) is not None

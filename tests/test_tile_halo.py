"""Halo of the sweep-block tiles: H rows, ceil(H/3) lozenge and ceil(H/2)
domino columns.

The sweep-block kernels load a tile with a halo of H rows but fewer
columns, and run a window of at most H pulling colour steps on it (a step
of colour c pulls the bits facing c from the colours rotated since c last
was; LozengeTiler.cpp:86-91 and DominoTiler.cpp:69-97 pick the colours,
lozengekernel.cl:38-53 and dominokernel*.cl UpdateTiles give the
neighbour offsets).  Checks that the tile interior stays exact, for the
lozenge model:

* the dependency frontier of the pull rule, over every colour sequence and
  every order of the colours' last rotations: after p pulls at most p rows
  and ceil(p / 3) columns of a plane are stale at a tile edge, and both
  bounds are reached;
* the reference's own step (the oracle's literal Rotate-then-Update walk):
  with garbage in every cell outside a region, a walk of a few steps
  leaves the region shrunk by q rows and ceil(q / 3) plane columns equal to
  the clean walk, for q the walk's pull count plus one (the reference's
  UpdateTiles after the last rotation spreads once more; the lazy engine
  makes that refresh once per walk, over the whole grid); shrinking by one
  plane column less fails for some seed, so the column bound is not slack;

and the same two for the domino model (every colour change is a pull; a
pull reads rows i-1, i+1 and column j - 1 or j + 1 by row parity): p rows
and ceil(p / 2) columns.
"""
import itertools
import math

import numpy as np

import randomtilingspublic_amd as rt
from tests import oracle_lib as o


def _pull_offsets(c):
    """Neighbour reads of a colour-c cell in row residue r (UpdateTiles with
    t = the source colour): pull A from a = c-1 reads (i-1, j+p2), (i, j-p4),
    (i+1, j); pull B from b = c-2 reads (i-1, j), (i, j+p8), (i+1, j-p16)."""
    a, b = (c + 2) % 3, (c + 1) % 3
    A, B = {}, {}
    for r in range(3):
        p2, p4 = int(r == (2 * a + 1) % 3), int(r == (2 * a + 2) % 3)
        p8, p16 = int(r == (2 * b) % 3), int(r == (2 * b + 1) % 3)
        A[r] = [(-1, p2), (0, -p4), (1, 0)]
        B[r] = [(-1, 0), (0, p8), (1, -p16)]
    return a, b, A, B


def _frontier(seq, order):
    """Stale columns at the right / left edge and stale rows at the top, per
    plane and row residue, after the colour sequence `seq`."""
    last = {order[0]: 0, order[1]: 1, order[2]: 2}
    g = 3
    R = [[0] * 3 for _ in range(3)]
    L = [[0] * 3 for _ in range(3)]
    T = [0, 0, 0]
    pulls = 0
    for c in seq:
        a, b, A, B = _pull_offsets(c)
        pa, pb = last[a] > last[c], last[b] > last[c]
        last[c] = g
        g += 1
        if not (pa or pb):
            continue
        pulls += 1
        nR, nL = list(R[c]), list(L[c])
        nT = T[c]
        for r in range(3):
            for src, on, offs in ((a, pa, A[r]), (b, pb, B[r])):
                if not on:
                    continue
                for di, dj in offs:
                    rr = (r + di) % 3
                    nR[r] = max(nR[r], R[src][rr] + dj)
                    nL[r] = max(nL[r], L[src][rr] - dj)
                nT = max(nT, T[src] + 1)
        R[c], L[c], T[c] = nR, nL, nT
    return pulls, max(map(max, R)), max(map(max, L)), max(T)


def test_pull_rule_frontier_rows_p_columns_ceil_p_over_3():
    worst = {}
    for n in range(1, 11):
        # only pulling steps matter (a step of the last colour moves nothing
        # and keeps the order), so colours never repeat back to back
        for first in range(3):
            for turns in itertools.product((1, 2), repeat=n - 1):
                seq = [first]
                for t in turns:
                    seq.append((seq[-1] + t) % 3)
                for order in itertools.permutations(range(3)):
                    p, right, left, top = _frontier(seq, order)
                    w = worst.setdefault(p, [0, 0, 0])
                    w[0], w[1], w[2] = max(w[0], right), max(w[1], left), max(w[2], top)
    for p, (right, left, top) in worst.items():
        assert top == p, (p, top)
        assert right == left == math.ceil(p / 3), (p, right, left)


def _pulls(seed, steps):
    cols = o.uniform012(seed, steps)
    last = [-1, -1, -1]
    pulls = 0
    for g, c in enumerate(cols):
        c = int(c)
        if last[(c + 2) % 3] > last[c] or last[(c + 1) % 3] > last[c]:
            pulls += 1
        last[c] = g
    return pulls


def test_reference_step_interior_survives_garbage_outside_the_tile():
    d = rt.LozengeTiler.AlmostHexDomain(40, 0)
    t0 = np.asarray(rt.LozengeTiler.MinTiling(d), dtype=np.int32)
    N = rt.side(t0)
    P = rt.synthetic_params(N * N, 91)
    t0 = o.lozenge_walk(t0, 3000, 5, P).reshape(N, N)  # a tiling with many flippable cells
    r0, r1, q0, q1 = 18, 62, 6, 21  # plane rows and plane columns of the "tile"
    rng = np.random.default_rng(3)
    slack_fails = 0
    for seed in range(1, 41):
        steps = 4 + seed % 9
        p = _pulls(seed, steps)
        bad = t0.copy()
        mask = np.ones((N, N), dtype=bool)
        mask[r0:r1, 3 * q0:3 * q1] = False
        bad[mask] = rng.integers(0, 64, size=int(mask.sum()))
        want = o.lozenge_walk(t0.ravel(), steps, seed, P).reshape(N, N)
        got = o.lozenge_walk(bad.ravel(), steps, seed, P).reshape(N, N)
        p += 1  # the final UpdateTiles
        h = math.ceil(p / 3)
        rows = slice(r0 + p, r1 - p)
        assert np.array_equal(got[rows, 3 * (q0 + h):3 * (q1 - h)],
                              want[rows, 3 * (q0 + h):3 * (q1 - h)]), (seed, steps, p)
        if h >= 1 and not np.array_equal(got[rows, 3 * (q0 + h - 1):3 * (q1 - h + 1)],
                                         want[rows, 3 * (q0 + h - 1):3 * (q1 - h + 1)]):
            slack_fails += 1
    assert slack_fails > 0


def _domino_frontier(p):
    """Stale columns at the right / left edge and stale rows at the top after
    p pulls, the colours alternating (UpdateTiles(tiling, reftiling, N, t):
    cell (i, j) of the other colour reads reftiling at (i-1, j), (i+1, j),
    (i, j - (i+t+1)%2) and (i, j + (i+t)%2))."""
    worst = (0, 0, 0)
    for first in (0, 1):
        R = [[0, 0], [0, 0]]
        L = [[0, 0], [0, 0]]
        T = [0, 0]
        c = first
        for _ in range(p):
            t = 1 - c  # the colour rotated last, pulled from
            nR, nL = list(R[c]), list(L[c])
            for r in range(2):
                right, left = (r + t) % 2, (r + t + 1) % 2
                for rr, dj in (((r - 1) % 2, 0), ((r + 1) % 2, 0), (r, -left), (r, right)):
                    nR[r] = max(nR[r], R[t][rr] + dj)
                    nL[r] = max(nL[r], L[t][rr] - dj)
            R[c], L[c], T[c] = nR, nL, max(T[c], T[t] + 1)
            c = t
        worst = tuple(max(a, b) for a, b in zip(worst, (max(map(max, R)), max(map(max, L)),
                                                         max(T))))
    return worst


def test_domino_pull_frontier_rows_p_columns_ceil_p_over_2():
    for p in range(1, 13):
        right, left, top = _domino_frontier(p)
        assert top == p
        assert right == left == math.ceil(p / 2), (p, right, left)


def _domino_pulls(seed, steps):
    cols = [int(x) % 2 for x in o.mt19937(seed, steps + 1)[1:]]  # the first output seeds TinyMT
    return sum(1 for a, b in zip(cols, cols[1:]) if a != b)


def test_domino_reference_step_interior_survives_garbage_outside_the_tile():
    t0 = np.asarray(rt.DominoTiler.MinTiling(rt.DominoTiler.AztecDiamond(40)), dtype=np.int32)
    N = rt.side(t0)
    P = rt.synthetic_params(N * N // 2, 92)
    t0 = o.domino_walk(t0, 4000, 5, P, rt.DOMINO_HEATBATH).reshape(N, N)
    r0, r1, q0, q1 = 20, 62, 10, 32  # rows and plane columns (2 tiling columns each)
    rng = np.random.default_rng(4)
    slack_fails = 0
    for seed in range(1, 41):
        steps = 4 + seed % 11
        p = _domino_pulls(seed, steps)
        bad = t0.copy()
        mask = np.ones((N, N), dtype=bool)
        mask[r0:r1, 2 * q0:2 * q1] = False
        bad[mask] = rng.integers(0, 16, size=int(mask.sum()))
        want = o.domino_walk(t0.ravel(), steps, seed, P, rt.DOMINO_HEATBATH).reshape(N, N)
        got = o.domino_walk(bad.ravel(), steps, seed, P, rt.DOMINO_HEATBATH).reshape(N, N)
        p += 1  # the final UpdateTiles
        h = math.ceil(p / 2)
        rows = slice(r0 + p, r1 - p)
        assert np.array_equal(got[rows, 2 * (q0 + h):2 * (q1 - h)],
                              want[rows, 2 * (q0 + h):2 * (q1 - h)]), (seed, steps, p)
        if h >= 1 and not np.array_equal(got[rows, 2 * (q0 + h - 1):2 * (q1 - h + 1)],
                                         want[rows, 2 * (q0 + h - 1):2 * (q1 - h + 1)]):
            slack_fails += 1
    assert slack_fails > 0

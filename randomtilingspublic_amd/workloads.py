"""Seeded synthetic domain families for batched runs (no reference
counterpart; the reference builds one domain per driver run, e.g.
`DominoTiler::AztecDiamond(n)` in Examples/AztecDiamondCftpEx.cpp:35).

`uneven_domino_batch(seed, n)` draws n domino domains: each is, with equal
odds, an Aztec diamond of order uniform in [32, 512], or a union of 1-8
axis-aligned rectangles with sides uniform in [16, 256], each new rectangle
placed over a cell of one already drawn (so the union is connected), holes
filled (so it is simply connected).  A union is redrawn while its bounding
box exceeds 512 cells, its area is odd, or Thurston's algorithm finds it
untileable.  Everything is drawn from numpy's PCG64 seeded with `seed`, so
a seed fixes the batch.
"""
import numpy as np

from . import DominoTiler, RTError


def rect_union_rects(rng, side=(16, 256), count=(1, 8), max_extent=512):
    """Rectangles (i0, j0, h, w) of one connected union, bounding box at
    most max_extent cells on each side."""
    while True:
        m = int(rng.integers(count[0], count[1] + 1))
        rects = []
        for k in range(m):
            h, w = (int(x) for x in rng.integers(side[0], side[1] + 1, size=2))
            if k == 0:
                i0 = j0 = 0
            else:
                a = rects[int(rng.integers(len(rects)))]
                pi = a[0] + int(rng.integers(a[2]))
                pj = a[1] + int(rng.integers(a[3]))
                i0 = pi - int(rng.integers(h))
                j0 = pj - int(rng.integers(w))
            rects.append((i0, j0, h, w))
        r = np.array(rects)
        ext = max((r[:, 0] + r[:, 2]).max() - r[:, 0].min(), (r[:, 1] + r[:, 3]).max() - r[:, 1].min())
        if ext <= max_extent:
            return rects


def tileable_rect_union(rng, **kw):
    """(rects, domain, min tiling) of the first drawn union that tiles."""
    while True:
        rects = rect_union_rects(rng, **kw)
        d = DominoTiler.RectangleUnion(rects)
        if int(d.sum()) % 2:
            continue
        try:
            return rects, d, DominoTiler.MinTiling(d)
        except RTError:
            continue


def uneven_domino_batch(seed, n=256, order=(32, 512), **kw):
    """List of n dicts {kind: "aztec" | "union", order | rects, domain,
    tiling (the minimal tiling), N (tiling side)}."""
    rng = np.random.default_rng(seed)
    aztec = {}
    out = []
    for _ in range(n):
        if int(rng.integers(2)) == 0:
            k = int(rng.integers(order[0], order[1] + 1))
            if k not in aztec:
                d = DominoTiler.AztecDiamond(k)
                aztec[k] = (d, DominoTiler.MinTiling(d))
            d, t = aztec[k]
            out.append({"kind": "aztec", "order": k, "domain": d, "tiling": t})
        else:
            rects, d, t = tileable_rect_union(rng, **kw)
            out.append({"kind": "union", "rects": rects, "domain": d, "tiling": t})
        out[-1]["N"] = int(round(len(out[-1]["tiling"]) ** 0.5))
    return out

"""Exact domino edge probabilities by Kasteleyn's method (FP64), used as the
statistical oracle for the samplers.

A domino tiling of a simply connected region is a perfect matching of its
squares; with K[b, w] = weight * (1 for a horizontal pair, i for a vertical
pair), |det K| is the weighted number of tilings and the probability of the
domino {b, w} is |K[b, w] * K^{-1}[w, b]|.

Position of a domino in the engine's vertex-state tilings (bits 1 up,
2 down, 4 left, 8 right on vertices; face (r, c) has corners (r, c) ..
(r+1, c+1)): the horizontal domino on faces (r, c), (r, c+1) is bit 2 of
vertex (r, c+1); the vertical domino on faces (r, c), (r+1, c) is bit 8 of
vertex (r+1, c).
"""
import numpy as np


def two_periodic_weights(a, b):
    """Per-domino weights whose products give the stationary law of the
    two-periodic rules (dominokerneltwoperiodic.cl:113-120): at vertex
    (i, j) the horizontal pair spanning columns j-1, j weighs a^2 if j is
    even else b^2, the vertical pair spanning rows i-1, i weighs a^2 if i
    is odd else b^2.  So a horizontal domino with left square in column c
    weighs a if c is odd, and a vertical domino with top square in row r
    weighs a if r is even."""
    def w(kind, r, c):
        if kind == "h":
            return a if c % 2 == 1 else b
        return a if r % 2 == 0 else b
    return w


def uniform_weights(kind, r, c):
    return 1.0


def edge_probabilities(domain, weight=uniform_weights):
    """domain: flat M x M 0/1 face array.  Returns (H, V): H[r, c] =
    P(horizontal domino on (r,c),(r,c+1)), V[r, c] = P(vertical domino on
    (r,c),(r+1,c)), both M x M."""
    d = np.asarray(domain).reshape(-1)
    M = int(round(np.sqrt(len(d))))
    d = d.reshape(M, M)
    faces = [(r, c) for r in range(M) for c in range(M) if d[r, c]]
    black = [f for f in faces if (f[0] + f[1]) % 2 == 0]
    white = [f for f in faces if (f[0] + f[1]) % 2 == 1]
    bi = {f: k for k, f in enumerate(black)}
    wi = {f: k for k, f in enumerate(white)}
    assert len(black) == len(white)
    K = np.zeros((len(black), len(white)), dtype=np.complex128)
    for (r, c), kb in bi.items():
        for dr, dc, ph in ((0, 1, 1.0), (0, -1, 1.0), (1, 0, 1j), (-1, 0, 1j)):
            nb = (r + dr, c + dc)
            if nb in wi:
                if dr == 0:
                    wgt = weight("h", r, min(c, c + dc))
                else:
                    wgt = weight("v", min(r, r + dr), c)
                K[kb, wi[nb]] = wgt * ph
    Kinv = np.linalg.inv(K)
    H = np.zeros((M, M))
    V = np.zeros((M, M))
    for (r, c), kb in bi.items():
        for dr, dc in ((0, 1), (0, -1), (1, 0), (-1, 0)):
            nb = (r + dr, c + dc)
            if nb not in wi:
                continue
            kw = wi[nb]
            p = abs(K[kb, kw] * Kinv[kw, kb])
            if dr == 0:
                H[r, min(c, c + dc)] = p
            else:
                V[min(r, r + dr), c] = p
    return H, V


def sample_edges(tilings, M):
    """Empirical counterpart of edge_probabilities for engine tilings
    (N = M + 1 vertex side)."""
    t = np.asarray(tilings).reshape(len(tilings), M + 1, M + 1)
    H = np.zeros((M, M))
    V = np.zeros((M, M))
    H[:, : M - 1] = ((t[:, :M, 1:M] & 2) != 0).mean(axis=0)
    V[: M - 1, :] = ((t[:, 1:M, :M] & 8) != 0).mean(axis=0)
    return H, V


def max_zscore(emp, exact, n):
    """Largest |emp - p| / sqrt(p(1-p)/n) over positions with 0 < p < 1,
    and positions where p is 0 or 1 must match exactly (returns inf)."""
    z = 0.0
    for e, p in ((emp[0], exact[0]), (emp[1], exact[1])):
        det = (p < 1e-12) | (p > 1 - 1e-12)
        if np.any(np.abs(e[det] - np.round(p[det])) > 0):
            return float("inf")
        m = ~det
        s = np.sqrt(p[m] * (1 - p[m]) / n)
        if m.any():
            z = max(z, float(np.max(np.abs(e[m] - p[m]) / s)))
    return z

"""Statistical pins of the oracle (and, with -m gpu, the device) against
mathematics it does not share code with: Kasteleyn FP64 edge probabilities
for uniform and two-periodic Aztec diamonds, and exact uniformity on tiny
regions whose tilings can be counted (Aztec order 2: 8 tilings; hexagon
2,2,2: MacMahon's 20)."""
import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import kasteleyn as kst
from tests import oracle_lib as o

Z_MAX = 5.0  # per-position bound; a few hundred positions -> false alarm < 1e-4


def _aztec(n):
    d = rt.DominoTiler.AztecDiamond(n)
    return d, rt.DominoTiler.MinTiling(d), rt.DominoTiler.MaxTiling(d)


def test_kasteleyn_counts_aztec():
    # |det K| of the Aztec diamond of order n is 2^(n(n+1)/2)
    for n in (1, 2, 3, 4):
        d, _, _ = _aztec(n)
        H, V = kst.edge_probabilities(d)
        M = int(round(np.sqrt(len(d))))
        # each square is covered exactly once
        dd = np.asarray(d).reshape(M, M)
        cover = np.zeros((M, M))
        cover += H
        cover[:, 1:] += H[:, :-1]
        cover += V
        cover[1:, :] += V[:-1, :]
        assert np.allclose(cover[dd == 1], 1.0)
    # corner dominoes of the Aztec diamond are forced? no: probability 1/2 by symmetry
    d, _, _ = _aztec(6)
    H, V = kst.edge_probabilities(d)
    assert abs(H.sum() + V.sum() - 6 * 7) < 1e-9  # n(n+1) dominoes


def test_oracle_cftp_matches_kasteleyn():
    d, tmin, tmax = _aztec(3)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, None)
    outs = []
    for m in range(1, 1201):
        out, _ = o.domino_cftp(tmax, tmin, m, P, 1, 64, 30, 1)
        outs.append(out)
    M = N - 1
    emp = kst.sample_edges(np.array(outs), M)
    exact = kst.edge_probabilities(d)
    assert kst.max_zscore(emp, exact, len(outs)) < Z_MAX


def test_oracle_cftp_uniform_on_aztec2():
    d, tmin, tmax = _aztec(2)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, None)
    seen = {}
    n = 1600
    for m in range(1, n + 1):
        out, _ = o.domino_cftp(tmax, tmin, m, P, 1, 32, 30, 1)
        seen[out.tobytes()] = seen.get(out.tobytes(), 0) + 1
    assert len(seen) == 8
    c = np.array(list(seen.values()), float)
    chi2 = ((c - n / 8) ** 2 / (n / 8)).sum()
    assert chi2 < 24.3  # chi^2_7 at p = 0.001


def test_oracle_lozenge_cftp_uniform_on_hexagon222():
    d = rt.LozengeTiler.AlmostHexDomain(2, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N, None)
    seen = {}
    n = 2000
    for m in range(1, n + 1):
        out, _ = o.lozenge_cftp(tmax, tmin, m, P, 1, 16, 30)
        seen[out.tobytes()] = seen.get(out.tobytes(), 0) + 1
    assert len(seen) == 20  # MacMahon: prod (i+j+k-1)/(i+j+k-2) = 20 for 2,2,2
    c = np.array(list(seen.values()), float)
    chi2 = ((c - n / 20) ** 2 / (n / 20)).sum()
    assert chi2 < 43.8  # chi^2_19 at p = 0.001


def test_oracle_walk_stationary_law_aztec2():
    """Walks from the min tiling, 300 steps each, 0.8-toggle rule: uniform."""
    d, tmin, _ = _aztec(2)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, None)
    seen = {}
    n = 1600
    for s in range(n):
        t = o.domino_walk(tmin, 300, 1000 + s, P, 0)
        seen[t.tobytes()] = seen.get(t.tobytes(), 0) + 1
    assert len(seen) == 8
    c = np.array(list(seen.values()), float)
    assert ((c - n / 8) ** 2 / (n / 8)).sum() < 24.3


@pytest.mark.gpu
@pytest.mark.parametrize("rule,a,b", [(rt.DOMINO_HEATBATH, 1.0, 1.0),
                                      (rt.DOMINO_HEATBATH2P, 1.0, 0.5)])
def test_device_cftp_matches_kasteleyn(gpu_ctx, rule, a, b):
    """Exact samples from the batched device CFTP on the Aztec diamond of
    order 10 against Kasteleyn FP64 probabilities (uniform and two-periodic
    a=1, b=0.5)."""
    d, tmin, tmax = _aztec(10)
    N = rt.side(tmin)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
    print("[case] device cftp kasteleyn", rule, a, b, flush=True)
    n = 600
    outs, _ = gpu_ctx.domino_cftp_batch(tmax, tmin, list(range(1, n + 1)), rt.SCHED_DOUBLING,
                                        256, 30, rule, a, b)
    M = N - 1
    emp = kst.sample_edges(outs, M)
    w = kst.uniform_weights if a == b else kst.two_periodic_weights(a, b)
    exact = kst.edge_probabilities(d, w)
    assert kst.max_zscore(emp, exact, n) < Z_MAX


@pytest.mark.parametrize("rule", [3, 2])
def test_oracle_two_periodic_matches_kasteleyn(rule):
    """Heat-bath CFTP (rule 3) and Metropolis walks (rule 2), a=1, b=0.5 on
    the Aztec diamond of order 3, against Kasteleyn with the two-periodic
    domino weights."""
    d, tmin, tmax = _aztec(3)
    N = rt.side(tmin)
    P = rt.synthetic_params(N * N // 2, None)
    outs = []
    n = 1500
    for m in range(1, n + 1):
        if rule == 3:
            out, _ = o.domino_cftp(tmax, tmin, m, P, 1, 64, 30, 3, 1.0, 0.5)
        else:
            out = o.domino_walk(tmin, 400, m, P, 2, 1.0, 0.5)
        outs.append(out)
    emp = kst.sample_edges(np.array(outs), N - 1)
    exact = kst.edge_probabilities(d, kst.two_periodic_weights(1.0, 0.5))
    assert kst.max_zscore(emp, exact, n) < Z_MAX
    # and the weighting is not a no-op: uniform probabilities are rejected
    assert kst.max_zscore(emp, kst.edge_probabilities(d), n) > Z_MAX

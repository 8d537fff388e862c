"""Serving: every request's sample equals the same request run alone (and
the oracle on small cases), whatever else shares the device with it."""
import time

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _domino(n):
    d = rt.DominoTiler.AztecDiamond(n)
    return rt.DominoTiler.MaxTiling(d), rt.DominoTiler.MinTiling(d)


def _lozenge(n, m=0):
    d = rt.LozengeTiler.AlmostHexDomain(n, m)
    return rt.LozengeTiler.MaxTiling(d), rt.LozengeTiler.MinTiling(d)


def _mix():
    reqs = []
    for k, n in enumerate([3, 6, 10, 5]):
        tmax, tmin = _domino(n)
        reqs.append(dict(model=rt.MODEL_DOMINO, tmax=tmax, tmin=tmin, master_seed=100 + k,
                         sched=rt.SCHED_DOUBLING, t0=128, rule=rt.DOMINO_HEATBATH))
    tmax, tmin = _domino(7)
    reqs.append(dict(model=rt.MODEL_DOMINO, tmax=tmax, tmin=tmin, master_seed=7,
                     sched=rt.SCHED_REFERENCE, t0=128, rule=rt.DOMINO_HEATBATH2P, a=1.0, b=0.5))
    for k, (n, m) in enumerate([(4, 0), (7, 2), (12, 0)]):
        tmax, tmin = _lozenge(n, m)
        reqs.append(dict(model=rt.MODEL_LOZENGE, tmax=tmax, tmin=tmin, master_seed=200 + k,
                         sched=rt.SCHED_DOUBLING, t0=64))
    return reqs


def _alone(ctx, q):
    if q["model"] == rt.MODEL_DOMINO:
        out, st = ctx.domino_cftp(q["tmax"], q["tmin"], q["master_seed"], q["sched"], q["t0"], 32,
                                  q["rule"], q.get("a", 1.0), q.get("b", 1.0))
    else:
        out, st = ctx.lozenge_cftp(q["tmax"], q["tmin"], q["master_seed"], q["sched"], q["t0"],
                                   32)
    return out, st


def test_served_equals_alone(gpu_ctx):
    reqs = _mix()
    big = max(rt.side(q["tmax"]) for q in reqs)
    P = rt.synthetic_params(big * big, 31)
    gpu_ctx.set_tinymt_params(P)
    alone = [_alone(gpu_ctx, q) for q in reqs]
    _say("server mix of", len(reqs))
    srv = rt.Server(gpu_ctx, 256 << 20, 64)
    try:
        ids = []
        for k, q in enumerate(reqs):
            ids.append(srv.submit(max_rounds=32, **q))
            if k % 3 == 2:
                time.sleep(0.01)  # later requests arrive while earlier ones run
        for rid, q, (want, wst) in zip(ids, reqs, alone):
            assert srv.wait(rid, 120.0)
            rc, got, st, lat = srv.poll(rid)
            assert rc == 0
            assert np.array_equal(got, want)
            assert st["rounds"] == wst["rounds"]
        s = srv.stats()
        assert s["completed"] == len(reqs) and s["status"] == 0
    finally:
        srv.close()
    # and the small domino ones against the oracle directly
    q = reqs[0]
    w, _ = o.domino_cftp(q["tmax"], q["tmin"], q["master_seed"], P, q["sched"], q["t0"], 32,
                         q["rule"])
    assert np.array_equal(alone[0][0], w)


def test_tight_budget_queues_and_cancel(gpu_ctx):
    reqs = _mix()[:4]
    big = max(rt.side(q["tmax"]) for q in reqs)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(big * big, 32))
    alone = [_alone(gpu_ctx, q) for q in reqs]
    # a budget that holds only one of these at a time
    tmax, _ = _domino(10)
    N = rt.side(tmax)
    # slot footprint: two TinyMT state sets, start + two sweep buffer sets of
    # both chains' planes (256-byte aligned each)
    one = 2 * (16 * (N - 2) * (N // 2 - 2) + 256) + 6 * (2 * N * (N // 2) + 256) + 4096
    _say("server tight budget", one)
    srv = rt.Server(gpu_ctx, one, 8)
    try:
        with pytest.raises(rt.RTError) as e:
            tm, tn = _domino(40)
            srv.submit(rt.MODEL_DOMINO, tm, tn, 1)
        assert e.value.code == -8  # larger than the whole budget
        ids = [srv.submit(max_rounds=32, **q) for q in reqs]
        victim = srv.submit(rt.MODEL_DOMINO, reqs[1]["tmax"], reqs[1]["tmin"], 999, t0=128)
        srv.cancel(victim)
        for rid, (want, _) in zip(ids, alone):
            assert srv.wait(rid, 120.0)
            rc, got, _, _ = srv.poll(rid)
            assert rc == 0 and np.array_equal(got, want)
        assert srv.wait(victim, 60.0)
        rc, got, _, _ = srv.poll(victim)
        assert rc in (-9, 0)  # cancelled unless it had already finished
    finally:
        srv.close()


def test_submitters_on_two_threads(gpu_ctx):
    """Requests submitted from two host threads while the loop admits and
    retires others: every result equals the request run alone, none is lost
    or admitted twice."""
    import threading
    base = _mix()
    reqs = []
    for rep in range(3):
        for q in base:
            reqs.append(dict(q, master_seed=q["master_seed"] + 1000 * rep))
    big = max(rt.side(q["tmax"]) for q in reqs)
    gpu_ctx.set_tinymt_params(rt.synthetic_params(big * big, 33))
    alone = [_alone(gpu_ctx, q)[0] for q in reqs]
    _say("server, two submitting threads,", len(reqs), "requests")
    srv = rt.Server(gpu_ctx, 64 << 20, 6)
    ids = [None] * len(reqs)
    errors = []

    def submit(part):
        try:
            for k in range(part, len(reqs), 2):
                ids[k] = srv.submit(max_rounds=32, **reqs[k])
                time.sleep(0.002 * (k % 3))
        except Exception as e:  # reported below
            errors.append(e)

    try:
        th = [threading.Thread(target=submit, args=(p,)) for p in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors
        assert len(set(ids)) == len(reqs)
        for rid, want in zip(ids, alone):
            assert srv.wait(rid, 120.0)
            rc, got, _, _ = srv.poll(rid)
            assert rc == 0 and np.array_equal(got, want)
        s = srv.stats()
        assert s["completed"] == len(reqs) and s["queued"] == 0 and s["status"] == 0
    finally:
        srv.close()


def test_walk_requests_in_the_mix(gpu_ctx):
    """Walk requests (RT_REQ_WALK) of all three models, rectangle-triangle
    included, served together with CFTP requests: each equals the model's
    plain Walk of the same start, steps and seed (rt_*_walk), and the small
    rectangle-triangle ones the oracle."""
    cftp = _mix()[:5]
    R = rt.RectTriangleTiler
    walks = [
        dict(model=rt.MODEL_DOMINO, start=_domino(12)[1], steps=700, seed=5,
             rule=rt.DOMINO_TOGGLE08),
        dict(model=rt.MODEL_DOMINO, start=_domino(9)[0], steps=333, seed=6,
             rule=rt.DOMINO_METRO2P, a=1.0, b=0.5),
        dict(model=rt.MODEL_LOZENGE, start=_lozenge(6, 1)[0], steps=501, seed=7),
        dict(model=rt.MODEL_RECTTRI, start=R.maxHex(3), steps=800, seed=8),
        dict(model=rt.MODEL_RECTTRI, start=R.minHex(5), steps=301, seed=9),
        dict(model=rt.MODEL_RECTTRI, start=R.maxHex(12), steps=1500, seed=10),
    ]
    big = max([rt.side(q["tmax"]) for q in cftp] + [rt.side(w["start"]) for w in walks])
    P = rt.synthetic_params(big * big, 34)
    gpu_ctx.set_tinymt_params(P)
    want_c = [_alone(gpu_ctx, q)[0] for q in cftp]
    want_w = []
    for w in walks:
        if w["model"] == rt.MODEL_DOMINO:
            want_w.append(gpu_ctx.domino_walk(w["start"], w["steps"], w["seed"], w["rule"],
                                              w.get("a", 1.0), w.get("b", 1.0)))
        elif w["model"] == rt.MODEL_LOZENGE:
            want_w.append(gpu_ctx.lozenge_walk(w["start"], w["steps"], w["seed"]))
        else:
            want_w.append(gpu_ctx.recttri_walk(w["start"], w["steps"], w["seed"]))
    _say("server: walks of three models among CFTP requests")
    srv = rt.Server(gpu_ctx, 256 << 20, 64)
    try:
        with pytest.raises(rt.RTError):  # rectangle-triangle has no CFTP
            srv.submit(rt.MODEL_RECTTRI, R.maxHex(3), R.minHex(3), 1)
        with pytest.raises(rt.RTError):
            srv.submit_walk(rt.MODEL_LOZENGE, walks[2]["start"], 0, 1)
        ids = []
        for k in range(max(len(cftp), len(walks))):
            if k < len(walks):
                w = walks[k]
                ids.append(("w", k, srv.submit_walk(w["model"], w["start"], w["steps"], w["seed"],
                                                    w.get("rule", rt.DOMINO_TOGGLE08),
                                                    w.get("a", 1.0), w.get("b", 1.0))))
            if k < len(cftp):
                ids.append(("c", k, srv.submit(max_rounds=32, **cftp[k])))
            time.sleep(0.005)
        for kind, k, rid in ids:
            assert srv.wait(rid, 120.0)
            rc, got, st, _ = srv.poll(rid)
            assert rc == 0, (kind, k)
            want = want_w[k] if kind == "w" else want_c[k]
            assert np.array_equal(got, want), (kind, k)
        assert srv.stats()["completed"] == len(ids)
    finally:
        srv.close()
    for k in (3, 4):
        w = walks[k]
        assert np.array_equal(want_w[k], o.recttri_walk(w["start"], w["steps"], w["seed"], P)), k


def test_sparse_domains_served_equal_alone(gpu_ctx):
    """Requests of which the server leaves tiles out (the cells that stay
    zero in every tiling): an L-shaped domino domain (most of its bounding
    square empty), an Aztec diamond and a hexagon with unequal sides, each
    equal to the same request run alone, and the L to the oracle."""
    D = rt.DominoTiler
    L = D.RectangleUnion([(0, 0, 110, 10), (100, 0, 10, 110)])
    reqs = [dict(model=rt.MODEL_DOMINO, tmax=D.MaxTiling(L), tmin=D.MinTiling(L), master_seed=11,
                 sched=rt.SCHED_DOUBLING, t0=128, rule=rt.DOMINO_HEATBATH)]
    tmax, tmin = _domino(40)
    reqs.append(dict(model=rt.MODEL_DOMINO, tmax=tmax, tmin=tmin, master_seed=12,
                     sched=rt.SCHED_DOUBLING, t0=128, rule=rt.DOMINO_HEATBATH))
    tmax, tmin = _lozenge(36, 12)
    reqs.append(dict(model=rt.MODEL_LOZENGE, tmax=tmax, tmin=tmin, master_seed=13,
                     sched=rt.SCHED_DOUBLING, t0=64))
    big = max(rt.side(q["tmax"]) for q in reqs)
    P = rt.synthetic_params(big * big, 32)
    gpu_ctx.set_tinymt_params(P)
    alone = [_alone(gpu_ctx, q) for q in reqs]
    _say("server: sparse domains")
    srv = rt.Server(gpu_ctx, 256 << 20, 64)
    try:
        ids = [srv.submit(max_rounds=32, **q) for q in reqs]
        for rid, (want, wst) in zip(ids, alone):
            assert srv.wait(rid, 120.0)
            rc, got, st, lat = srv.poll(rid)
            assert rc == 0 and np.array_equal(got, want) and st["rounds"] == wst["rounds"]
    finally:
        srv.close()
    q = reqs[0]
    w, _ = o.domino_cftp(q["tmax"], q["tmin"], q["master_seed"], P, q["sched"], q["t0"], 32,
                         q["rule"])
    assert np.array_equal(alone[0][0], w)

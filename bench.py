#!/usr/bin/env python3
"""Benchmark of the random-tiling engine on one MI355X.

Default workload (the headline configuration): uniform lozenge tilings of
the hexagon with sides a=b=c=300, exact samples by coupling from the past
(top = MaxTiling, bottom = MinTiling, look-back doubling from 2^10 steps),
one sample per master seed, seeds 1..32, all 32 samples advanced together
by the batched fused kernel.  A "step" of this benchmark is one exact
sample (one master seed).  The timed batch is always the full
configuration, seeds 1..32: --steps does not shrink it (the line reports
the 32 samples it timed as `steps` and the requested value as
`steps_requested`); --samples changes it for profiling runs only.
--warmup W short walks run first.  Synthetic TinyMT parameter table (the
TINYMT32J set for every stream; the reference's 2^20-line table is not in
its checkout).

Prints ONE JSON line: value = CFTP samples/s; also flip attempts/s (one
attempt = one chain's flip attempt at one site, as the reference counts
them: each CFTP round is two reference Walks), the plain one-chain-per-call
path on the same inputs, the roofline of the sweep kernel and the CPU
baseline (the oracle, OpenMP over the sites, on this host's cores).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
XCD_ORDER = None  # --xcd-order: rt_set_xcd_order of the measured contexts (None: the library's)
STREAMS = 0  # --streams: rt_set_streams of the headline and batch-walk contexts (0: the library's)
# Sweep-block kernel instantiation per launch-shape slot (rt_timing.shape_*[k]):
# k_fused<model 1 = lozenge, chains per workgroup, waves, rows per lane>; slots
# 4..6 run CFTP pairs split, one chain per workgroup; slot 7 is the 48-row
# pair kernel at eight waves per SIMD.
KERNEL_BY_SHAPE = ["k_fused<1, 2, 8, 8>", "k_fused<1, 2, 16, 4>", "k_fused<1, 2, 16, 2>",
                   "k_fused<2, 2, 8, 4>", "k_fused<1, 1, 8, 8>", "k_fused<1, 1, 16, 4>",
                   "k_fused<1, 1, 16, 2>", "k_pair_lozenge<32>"]


def set_halo_kernels(halo):
    """Slot 7 runs k_pair_lozenge (both chains packed per LDS word) when the
    halo is the default (0: 9 for these batches) or a multiple of 3, else
    k_fused_o8<1>."""
    KERNEL_BY_SHAPE[7] = "k_pair_lozenge<32>" if halo % 3 == 0 else "k_fused_o8<1>"
# Algorithmic bytes per site per launch: a pair workgroup reads the TinyMT
# state (16 B) and parameters (12 B) once, writes the state (16 B), reads and
# writes 3 colour planes x 2 chains (12 B) = 56 B; split pairs read state and
# parameters in both workgroups (2 x 28 B), write the state once (16 B) and
# the planes as before (12 B) = 84 B.
ALG_BYTES_BY_SHAPE = [56, 56, 56, 56, 84, 84, 84, 56]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tmax, tmin, params, units_attempts, nsamples, threads, t0=1024, rounds=6):
    """The oracle (CPU restatement of the reference kernels and CFTP driver,
    OpenMP over each NDRange) timed on a bounded sample of the same
    workload: the CFTP of master seed 1 on the same hexagon, stopped after
    `rounds` doubling rounds (both chains, every segment, as the device
    runs them; 2 x 64512 steps of 123627 sites at the default, ~10-30 s on
    16 cores).  Scaled to samples/s with the flip attempts per sample the
    device run needed."""
    from tests import oracle_lib as o
    import ctypes as C
    o.set_threads(threads)
    N = int(round(np.sqrt(len(tmin))))
    P = N * (N // 3)
    out = np.zeros_like(tmin)
    st = o.CftpStats()
    p = o._p(params)
    t1 = time.time()
    rc = o._L.orc_lozenge_cftp(tmax, tmin, N, 1, 1, t0, rounds, p, len(p) // 3, out,
                               C.byref(st))
    dt = time.time() - t1
    o.set_threads(1)
    steps = sum(t0 << max(r - 1, 0) for r in range(1, rounds + 1))  # doubling look-backs
    done = 2 * steps * P
    att_s = done / dt
    return {
        "value": att_s * nsamples / units_attempts if units_attempts else None,
        "unit": "CFTP samples/s",
        "cores": threads,
        "kind": "port",
        "attempts_per_s": att_s,
        "sample": f"oracle CFTP of master seed 1 on the same hexagon, {rounds} doubling rounds "
                  f"from 2^{int(np.log2(t0))} ({steps} steps x 2 chains, "
                  f"{'coalesced' if rc == 0 else 'not coalesced, as expected'}) in {dt:.1f} s, "
                  "scaled by the device run's flip attempts per sample",
    }


def plain_sweep(rt, order=3000, steps=1000):
    """The plain Aztec-diamond sweep at HBM scale: heat-bath domino walk on
    the Aztec diamond of order 3000 (N = 6006, 18.0 M TinyMT streams: 505 MB
    of state and parameters plus 36 MB of colour planes, twice the 256 MB
    last-level cache), one launch per colour step (the reference's launch
    shape, ENGINE_STEP).  Algorithmic bytes per flip attempt: TinyMT state
    16 B read + 16 B written + parameters 12 B, the rotated cell 1 B read +
    1 B written and its four neighbours 1 B each read once per step = 48 B."""
    d = rt.DominoTiler.AztecDiamond(order)
    t = rt.DominoTiler.MinTiling(d)
    N = rt.side(t)
    ctx = rt.Context(0)
    try:
        ctx.set_engine(rt.ENGINE_STEP)
        ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
        ctx.domino_walk(t, 50, 1, rt.DOMINO_HEATBATH)
        ctx.domino_walk(t, steps, 1, rt.DOMINO_HEATBATH)
        tm = ctx.last_timing()
    finally:
        ctx.close()
    att_s = tm["attempts"] / (tm["device_ms"] / 1e3)
    gbs = 48.0 * att_s / 1e9
    return {"workload": f"domino Aztec diamond order {order} (N={N}), heat-bath, {steps} steps, "
                        "one launch per colour step",
            "kernel": "k_domino_step<1>", "us_per_step": round(tm["device_ms"] * 1e3 / steps, 2),
            "sites_per_launch": (N - 2) * (N // 2 - 2),
            "avg_launch_us": round(tm["device_ms"] * 1e3 / steps, 2),
            "flip_attempts_per_s": round(att_s, 1),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "alg_bytes_per_attempt": 48}}


def batched_walks(rt, seed=4, nchains=256, steps=10 ** 4):
    """Batched configuration: 256 independent heat-bath domino chains on
    uneven domains drawn from `seed` (Aztec diamonds of order U[32, 512]
    and tileable unions of 1-8 rectangles with sides U[16, 256],
    randomtilingspublic_amd.workloads), each started from its minimal
    tiling and walked 10^4 steps, all chains advanced by the same sweep-block
    launches, the launch sequence captured in one hipGraph and replayed; the
    tile shape comes from the autotuner (tuned on a 400-step warm-up batch).
    Algorithmic bytes per site per launch of k_fused<0, 1, *, *>: TinyMT
    state 16 B read + 16 B written + parameters 12 B + the two colour planes
    2 B read + 2 B written = 48 B."""
    from randomtilingspublic_amd import workloads as W
    t0 = time.time()
    batch = W.uneven_domino_batch(seed, nchains)
    gen_s = time.time() - t0
    tilings = [b["tiling"] for b in batch]
    big = max(b["N"] for b in batch)
    sites = sum((b["N"] - 2) * (b["N"] // 2 - 2) for b in batch)
    ctx = rt.Context(0)
    if XCD_ORDER is not None:
        ctx.set_xcd_order(XCD_ORDER)
    if STREAMS:
        ctx.set_streams(STREAMS)
    try:
        ctx.set_tinymt_params(rt.synthetic_params(big * big // 2, None))
        seeds = [1 + k for k in range(nchains)]
        ctx.domino_walk_batch(tilings, 400, seeds, rt.DOMINO_HEATBATH, use_graph=False)
        warm = ctx.last_timing()
        w = time.time()
        ctx.domino_walk_batch(tilings, steps, seeds, rt.DOMINO_HEATBATH, use_graph=True)
        wall = time.time() - w
        tm = ctx.last_timing()
    finally:
        ctx.close()
    dev_s = tm["device_ms"] / 1e3
    att_s = tm["attempts"] / dev_s
    launches = tm["launches"]
    alg = 48.0 * tm["kernel_sites"]  # sites of the chains each launch advances
    gbs = alg / dev_s / 1e9
    foot = sum(b["N"] * b["N"] // 2 * (16 + 2) for b in batch)
    # VALU issue over the device time: wave instructions per site per launch
    # of the packed kernel (committed SQ profile of this workload) x sites
    valu = None
    sqfile = os.path.join(ROOT, "profiles", "kernels_sq.json")
    if os.path.exists(sqfile) and tm["shape"] == 7:
        with open(sqfile) as f:
            sq = json.load(f).get("k_pair_domino")
        if sq and sq["workload"] == "batch256_walk":
            ins = sq["valu_wave_insts_per_site"] * tm["kernel_sites"]
            valu = {"issue_frac": round(ins / (SIMD_WAVE_INSTS_PER_US * dev_s * 1e6), 4),
                    "source": "profiles/kernels_sq.json (" + sq["profile"] + ")"}
    return {"workload": f"{nchains} heat-bath domino chains on uneven domains from seed {seed} "
                        f"(Aztec order U[32,512] and rectangle unions), {steps} steps each, "
                        "one hipGraph, autotuned tile shape",
            "chains": nchains, "aztec": sum(b["kind"] == "aztec" for b in batch),
            "N_min": min(b["N"] for b in batch), "N_max": big, "sites_per_step": sites,
            "state_footprint_MB": round(foot / 1e6, 1), "generate_s": round(gen_s, 2),
            "wall_s": round(wall, 3), "device_s": round(dev_s, 3),
            "flip_attempts_per_s": round(att_s, 1),
            "tile_shape": tm["shape"], "tuned_on_warmup": warm["shape_source"] == 1,
            "launches": launches,
            "sites_per_launch": int(tm["kernel_sites"] / max(tm["kernel_launches"], 1)),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "alg_bytes_per_site_per_launch": 48,
                         "note": "device time includes the final colour-plane sync",
                         "valu": valu}}


def weighted_and_bibone(rt, cftp_seeds=8, cftp_order=24, steps=10 ** 5, rounds512=7):
    """Weighted 2-periodic domino walk on the Aztec diamond of order 512
    (a = 1.0, b = 0.5, the reference's Metropolis rule of
    dominokerneltwoperiodic.cl, 10^5 steps from the minimal tiling), the
    bibone walk on the hexagon with sides 200, 201, 200 (10^5 steps), and
    the two-periodic heat-bath CFTP (dominokernelCFTPtwoperiodic.cl, same
    weights) for master seeds 1..cftp_seeds, batched: to completion on the
    Aztec diamond of order `cftp_order`, and on order 512 for `rounds512`
    doubling rounds (look-back 2^10 .. 2^(9+rounds512)), reporting after
    each round how many vertex states still differ between the coupled
    chains.  That chain coalesces far more slowly than the uniform one
    (order 24: 2^16 steps against 2^13; order 64: not within a 2^22-step
    look-back), so order 512 does not coalesce within reach."""
    D = rt.DominoTiler
    out = {}
    d = D.AztecDiamond(512)
    tmin = D.MinTiling(d)
    N = rt.side(tmin)
    ctx = rt.Context(0)
    try:
        ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
        ctx.domino_walk(tmin, 200, 1, rt.DOMINO_METRO2P, 1.0, 0.5)
        w = time.time()
        ctx.domino_walk(tmin, steps, 1, rt.DOMINO_METRO2P, 1.0, 0.5)
        wall = time.time() - w
        tm = ctx.last_timing()
        out["weighted_walk"] = {
            "workload": f"domino Aztec diamond order 512 (N={N}), two-periodic Metropolis "
                        f"a=1.0 b=0.5, {steps} steps from the minimal tiling",
            "wall_s": round(wall, 3), "device_s": round(tm["device_ms"] / 1e3, 3),
            "us_per_step": round(tm["device_ms"] * 1e3 / steps, 3),
            "flip_attempts_per_s": round(tm["attempts"] / (tm["device_ms"] / 1e3), 1),
            "launches": tm["launches"], "tile_shape": tm["shape"]}
        if cftp_seeds:
            dc = D.AztecDiamond(cftp_order)
            cmin, cmax = D.MinTiling(dc), D.MaxTiling(dc)
            w = time.time()
            outs, stats = ctx.domino_cftp_batch(cmax, cmin, list(range(1, cftp_seeds + 1)),
                                                rt.SCHED_DOUBLING, 1024, 16,
                                                rt.DOMINO_HEATBATH2P, 1.0, 0.5)
            wall = time.time() - w
            out["weighted_cftp"] = {
                "workload": f"two-periodic heat-bath CFTP (a=1.0, b=0.5) of the Aztec diamond "
                            f"of order {cftp_order}, doubling from 2^10, master seeds "
                            f"1..{cftp_seeds}, batched",
                "wall_s": round(wall, 2), "samples_per_s": round(cftp_seeds / wall, 4),
                "rounds": [s["rounds"] for s in stats]}
        if cftp_seeds and rounds512:
            # tmin / d are the order-512 diamond of the walk above
            tmax = D.MaxTiling(d)
            masters = list(range(1, cftp_seeds + 1))
            differ, dev_ms, att = [], 0.0, 0
            for r in range(1, rounds512 + 1):
                tops, bots, st = ctx.domino_cftp_batch_rounds(
                    tmax, tmin, masters, rt.SCHED_DOUBLING, 1024, r, rt.DOMINO_HEATBATH2P,
                    1.0, 0.5)
                differ.append(int(np.count_nonzero(tops != bots)))
                tm = ctx.last_timing()
                dev_ms += tm["device_ms"]
                att += tm["attempts"]
            out["weighted_cftp_512"] = {
                "workload": f"two-periodic heat-bath CFTP (a=1.0, b=0.5) of the Aztec diamond "
                            f"of order 512 (N={N}), master seeds 1..{cftp_seeds}, batched, "
                            f"doubling from 2^10, rounds 1..{rounds512} (each call reruns "
                            "rounds 1..r)",
                "differing_vertices_after_round": differ,
                "vertices": cftp_seeds * N * N,
                "coalesced": differ[-1] == 0,
                "device_s": round(dev_ms / 1e3, 3),
                "flip_attempts_per_s": round(att / (dev_ms / 1e3), 1)}
        out["bibone_walk"] = bibone_walk(rt, ctx, steps)
    finally:
        ctx.close()
    return out


def two_periodic_scan(rt, orders=(16, 20, 24, 28, 32), seeds=8, max_rounds=15):
    """Coalescence of the two-periodic heat-bath CFTP (a = 1.0, b = 0.5,
    dominokernelCFTPtwoperiodic.cl) against the Aztec order: master seeds
    1..seeds batched, doubling from 2^10, at most max_rounds rounds
    (look-back up to 2^(9 + max_rounds)); per order the rounds each seed
    needed (max_rounds + 1: not within the cap) and the device time."""
    D = rt.DominoTiler
    out = {"workload": f"two-periodic heat-bath CFTP (a=1.0, b=0.5), Aztec orders {list(orders)}, "
                       f"master seeds 1..{seeds}, doubling from 2^10, at most {max_rounds} "
                       "rounds", "orders": {}}
    ctx = rt.Context(0)
    ctx.set_round_callback(lambda r, n, fin, tot: print(
        f"[bench] two-periodic round {r} ({n} steps): {fin} of {tot} coalesced", file=sys.stderr,
        flush=True))
    try:
        for n in orders:
            d = D.AztecDiamond(n)
            tmin, tmax = D.MinTiling(d), D.MaxTiling(d)
            N = rt.side(tmin)
            ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
            masters = list(range(1, seeds + 1))
            t1 = time.time()
            _, stats = ctx.domino_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, 1024,
                                             max_rounds, rt.DOMINO_HEATBATH2P, 1.0, 0.5,
                                             partial=True)
            rounds = [st["rounds"] if st["total_steps"] else max_rounds + 1 for st in stats]
            wall = time.time() - t1
            out["orders"][n] = {"rounds": rounds, "wall_s": round(wall, 2),
                                "median_lookback_log2": float(np.median(rounds)) + 9}
            print(f"[bench] two-periodic order {n}: rounds {rounds} in {wall:.1f} s",
                  file=sys.stderr, flush=True)
    finally:
        ctx.close()
    return out


def bibone_walk(rt, ctx, steps=10 ** 5):
    """The bibone walk on the hexagon with sides 200, 201, 200 (tiling side
    812, 406 x 406 plane cells), `steps` steps from the hexagon's tiling, on
    the edge form (k_bi_edge).  Algorithmic bytes per plane cell per launch:
    the edge byte read and written (2 B), TinyMT state read and written
    (32 B) and parameters read (12 B) = 46 B."""
    B = rt.TriangleDimerTiler
    t = B.Hexagon(200, 200, 201)
    Nb = rt.side(t)
    M = Nb // 2
    ctx.set_tinymt_params(rt.synthetic_params(M * M, None))
    ctx.bibone_walk(t, 200, 1)
    w = time.time()
    ctx.bibone_walk(t, steps, 1)
    wall = time.time() - w
    tm = ctx.last_timing()
    return {"workload": f"bibone hexagon sides 200, 201, 200 (tiling side {Nb}), {steps} steps",
            "wall_s": round(wall, 3), "device_s": round(tm["device_ms"] / 1e3, 3),
            "us_per_step": round(tm["device_ms"] * 1e3 / steps, 3),
            "flip_attempts_per_s": round(tm["attempts"] / (tm["device_ms"] / 1e3), 1),
            "launches": tm["launches"], "tile_shape": tm["shape"], "sites_per_launch": M * M,
            "avg_launch_us": round(tm["device_ms"] * 1e3 / tm["launches"], 2)}


def bibone_section(rt, shape=0, steps=10 ** 5):
    ctx = rt.Context(0)
    try:
        ctx.set_tile_shape(shape)
        if XCD_ORDER is not None:
            ctx.set_xcd_order(XCD_ORDER)
        return bibone_walk(rt, ctx, steps)
    finally:
        ctx.close()


def served_stream(rt, seconds=40.0, seed=2024, dom=(64, 256), loz=(64, 160), t0=1024, burst_n=32,
                  plain_n=6, load=0.8, rtri=(32, 256), rtri_steps=10 ** 4):
    """A Poisson stream of CFTP requests through the continuous-batching
    server (rt_server), then the same requests one call at a time.  Request
    mix, drawn from `seed`: domino Aztec diamonds of order dom[0]..dom[1]
    (heat bath) and lozenge hexagons of side loz[0]..loz[1], doubling from
    t0, and rectangle-triangle hexagons of side rtri[0]..rtri[1] as walks of
    rtri_steps steps from minHex (that model has no CFTP: its coupling is
    not monotone; the reference's own example walks from minHex), each
    model with probability 1/3 (1/2 each without rtri).  The
    arrival rate is `load` times the capacity measured on a burst of burst_n
    requests (the served rate follows the offered one; latency is the
    price of a higher load).  Reports samples/s, p50/p99 latency (submit to result) and
    the plain path's samples/s on the first plain_n requests."""
    rng = np.random.default_rng(seed)
    cache = {}

    def domain(model, n):
        if (model, n) not in cache:
            if model == rt.MODEL_DOMINO:
                d = rt.DominoTiler.AztecDiamond(n)
                cache[(model, n)] = (rt.DominoTiler.MaxTiling(d), rt.DominoTiler.MinTiling(d))
            elif model == rt.MODEL_RECTTRI:
                cache[(model, n)] = (rt.RectTriangleTiler.minHex(n), None)
            else:
                d = rt.LozengeTiler.AlmostHexDomain(n, 0)
                cache[(model, n)] = (rt.LozengeTiler.MaxTiling(d), rt.LozengeTiler.MinTiling(d))
        return cache[(model, n)]

    sizes = {rt.MODEL_DOMINO: dom, rt.MODEL_LOZENGE: loz}
    if rtri:
        sizes[rt.MODEL_RECTTRI] = rtri

    def draw(k):
        model = int(rng.integers(0, len(sizes)))
        lo, hi = sizes[model]
        n = int(rng.integers(lo, hi + 1))
        tmax, tmin = domain(model, n)
        return dict(model=model, n=n, tmax=tmax, tmin=tmin, master_seed=10000 + k)

    big = max(rt.side(domain(m, hi)[0]) for m, (lo, hi) in sizes.items())
    ctx = rt.Context(0)
    ctx.set_tinymt_params(rt.synthetic_params(big * big, None))
    srv = rt.Server(ctx, 8 << 30, max_active=256)
    try:
        def submit(q):
            if q["model"] == rt.MODEL_RECTTRI:
                return srv.submit_walk(q["model"], q["tmax"], rtri_steps, q["master_seed"])
            return srv.submit(q["model"], q["tmax"], q["tmin"], q["master_seed"],
                              rt.SCHED_DOUBLING, t0, 32, rt.DOMINO_HEATBATH)

        def finish(i, what):
            """Poll request i until it is done (a progress line per minute)."""
            t1, last = time.time(), time.time()
            while True:
                r = srv.poll(i)
                if r is not None:
                    assert r[0] == 0, r[0]
                    return r
                srv.wait(i, 1.0)
                if time.time() - last > 60:
                    last = time.time()
                    print(f"[bench] served: {what} still running after {time.time() - t1:.0f} s",
                          file=sys.stderr, flush=True)

        # capacity on a burst
        burst = [draw(k) for k in range(burst_n)]
        tb = time.time()
        ids = [submit(q) for q in burst]
        for n, i in enumerate(ids):
            finish(i, f"burst request {n}")
        cap = len(burst) / (time.time() - tb)
        rate = load * cap
        print(f"[bench] served: burst capacity {cap:.1f} samples/s, stream at {rate:.1f}/s",
              file=sys.stderr, flush=True)
        # Poisson stream
        stream, lat, pending = [], [], {}
        start = time.time()
        nxt = start
        k = 1000
        last_report = start
        while time.time() - start < seconds:
            now = time.time()
            while nxt <= now:
                q = draw(k)
                k += 1
                pending[submit(q)] = q
                stream.append(q)
                print(f"[bench] served: request {len(stream)} arrives (model {q['model']}, "
                      f"size {q['n']})", file=sys.stderr, flush=True)
                nxt += rng.exponential(1.0 / rate)
            for i in list(pending):
                r = srv.poll(i)
                if r is not None:
                    assert r[0] == 0
                    lat.append((pending[i]["model"], r[3]))
                    del pending[i]
                    print(f"[bench] served: a request completed after {r[3]:.1f} s",
                          file=sys.stderr, flush=True)
            if now - last_report > 60:
                last_report = now
                st = srv.stats()
                print(f"[bench] served: {st['completed']} completed, {len(pending)} in flight, "
                      f"{st['ticks']} device ticks", file=sys.stderr, flush=True)
            time.sleep(0.002)
        for n, i in enumerate(list(pending)):
            r = finish(i, f"stream request {n} of {len(pending)} left")
            lat.append((pending[i]["model"], r[3]))
            print(f"[bench] served: drained {n + 1} of {len(pending)}", file=sys.stderr,
                  flush=True)
        span = time.time() - start
    finally:
        srv.close()
    served = len(stream) / span
    # The same requests one call at a time (the first plain_n of them,
    # bounded): with the reference's launch shape (one launch per colour
    # step) and with the sweep-block kernel.
    def plain_rate(engine, reqs):
        ctx.set_engine(engine)
        t1 = time.time()
        for n, q in enumerate(reqs):
            print(f"[bench] served: plain request {n} (model {q['model']}, size {q['n']}), "
                  f"engine {engine}", file=sys.stderr, flush=True)
            if q["model"] == rt.MODEL_DOMINO:
                ctx.domino_cftp(q["tmax"], q["tmin"], q["master_seed"], rt.SCHED_DOUBLING, t0,
                                32, rt.DOMINO_HEATBATH)
            elif q["model"] == rt.MODEL_RECTTRI:
                ctx.recttri_walk(q["tmax"], rtri_steps, q["master_seed"])
            else:
                ctx.lozenge_cftp(q["tmax"], q["tmin"], q["master_seed"], rt.SCHED_DOUBLING, t0,
                                 32)
        return len(reqs) / (time.time() - t1)

    sub = stream[:plain_n]
    plain = plain_rate(rt.ENGINE_STEP, sub) if sub else None
    plain_fused = plain_rate(0, sub) if sub else None
    ctx.set_engine(0)
    ctx.close()
    lat_cftp = np.array([v for m, v in lat if m != rt.MODEL_RECTTRI])
    lat = np.array([v for _, v in lat])
    mix = (f"CFTP requests of domino Aztec order {dom[0]}-{dom[1]} (heat bath) and lozenge "
           f"hexagon side {loz[0]}-{loz[1]}, doubling from 2^{int(np.log2(t0))}")
    if rtri:
        mix += (f", and walks of {rtri_steps} steps from minHex of rectangle-triangle hexagon "
                f"side {rtri[0]}-{rtri[1]} (no CFTP for that model), 1/3 each")
    return {"workload": f"Poisson stream of {mix}; rate {round(100 * load)} % of burst "
                        "capacity, continuous batching",
            "requests": len(stream), "span_s": round(span, 2),
            "samples_per_s": round(served, 4), "burst_capacity_per_s": round(cap, 4),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "cftp_requests": int(lat_cftp.size),
            "p50_cftp_latency_s": round(float(np.percentile(lat_cftp, 50)), 4)
            if lat_cftp.size else None,
            "p99_cftp_latency_s": round(float(np.percentile(lat_cftp, 99)), 4)
            if lat_cftp.size else None,
            "plain_samples_per_s": round(plain, 4) if plain else None,
            "plain_fused_samples_per_s": round(plain_fused, 4) if plain_fused else None,
            "plain_requests": len(sub),
            "served_vs_plain": round(served / plain, 2) if plain else None,
            "capacity_vs_plain": round(cap / plain, 2) if plain else None,
            "capacity_vs_plain_fused": round(cap / plain_fused, 2) if plain_fused else None}


def copy_probe(mib):
    """Plain device-to-device copy of `mib` MiB (hipMemcpy, which runs the
    runtime's copy kernel), five times: a kernel whose HBM bytes are known
    (mib MiB read, mib MiB written per launch), to scale the FETCH_SIZE /
    WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 0x5A, C.c_size_t(n)) == 0
    ev = [C.c_void_p(), C.c_void_p()]
    for e in ev:
        assert hip.hipEventCreate(C.byref(e)) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipEventRecord(ev[0], None)
    for _ in range(5):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipEventRecord(ev[1], None)
    assert hip.hipEventSynchronize(ev[1]) == 0
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1])
    ms = ms.value / 5
    probe = (C.c_ubyte * 4096)()
    assert hip.hipMemcpy(probe, C.c_void_p(dst.value + n - 4096), C.c_size_t(4096), 2) == 0
    assert all(b == 0x5A for b in probe)
    for e in ev:
        hip.hipEventDestroy(e)
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n, "bytes_written": n,
                      "avg_ms": round(ms, 4), "GBs": round(2 * n / ms / 1e6, 1)}), flush=True)


# Profiling workloads of --rocprof-profile: name -> (bench.py arguments,
# kernel whose dispatches the summary is about, algorithmic bytes per site
# per dispatch of that kernel).
PROFILE_WORKLOADS = {
    "hex300_cftp_32seeds": (["--profile-rounds", "6", "--samples", "32", "--tile-shape", "7",
                             "--streams", "1"], "k_pair_lozenge<32>", 56),
    "aztec3000_step": (["--profile-section", "sweep"], "k_domino_step<1>", 48),
    "batch256_walk": (["--profile-section", "walks", "--streams", "1"], "k_pair_domino", 48),
    "bibone_hex200": (["--profile-section", "bibone", "--bibone-steps", "20000"],
                      "k_bi_edge<64, 16, 1024>", 46),
}
PROFILE_TEXT = {
    "hex300_cftp_32seeds": "lozenge hexagon a=b=c=300 (N=609, 123 627 sites per plane), batched "
                           "CFTP of master seeds 1..32, doubling from 2^10, stopped after 6 "
                           "doubling rounds, tile shape 7 (k_pair_lozenge: 96 x 32 chain-pair "
                           "tiles, halo 9 rows and 3 columns, eight waves per SIMD), one "
                           "stream, so one dispatch per launch window covers every sample (the "
                           "benchmark splits them over two); 56 B per "
                           "site: TinyMT state 16 B read + 16 B written, "
                           "parameters 12 B, 3 colour planes x 2 chains read and written",
    "aztec3000_step": "plain heat-bath domino sweep of the Aztec diamond of order 3000 "
                      "(N=6006), 1000 steps, one launch per colour step; 48 B per attempt: "
                      "TinyMT state 16 + 16 B, parameters 12 B, cells 4 B",
    "batch256_walk": "256 heat-bath domino chains on uneven domains (seed 4), 10^4 steps, one "
                     "hipGraph (one stream here; the benchmark's graph has two parallel "
                     "chains of launches), tile shape 7 (k_pair_domino, "
                     "one chain in the low halves, 96 x 32 tiles, halo 8 rows and 4 columns); "
                     "48 B per site per launch: TinyMT state 16 + 16 B, parameters 12 B, 2 "
                     "colour planes read and written",
    "bibone_hex200": "bibone walk on the hexagon with sides 200, 201, 200 (tiling side 812, "
                     "406 x 406 plane cells), 2 x 10^4 steps on the edge form (one bit per "
                     "lattice edge), 64 x 64-cell tiles with halo 16, 8 steps per launch; 46 B "
                     "per plane cell per launch: the edge byte read and written, TinyMT state "
                     "16 + 16 B, parameters 12 B",
}


def _short_kernel(name):
    """'void rt::k_fused_o8<1>(rt::UnitDesc const*, ...)' -> 'k_fused_o8<1>'."""
    n = name.strip().strip('"').replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):  # drop the argument list (outside template brackets)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    n = n[:cut]
    return n.replace("rt::", "").replace("(anonymous namespace)::", "").strip()


def _csv_rows(root, suffix):
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(root, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def _kernel_stats(root):
    out = {}
    for r in _csv_rows(root, "kernel_stats.csv"):
        k = _short_kernel(r["Name"])
        out[k] = {"calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) / 1e6,
                  "avg_us": float(r["AverageNs"]) / 1e3, "min_us": float(r["MinNs"]) / 1e3,
                  "max_us": float(r["MaxNs"]) / 1e3, "pct": float(r["Percentage"])}
    return out


def _counter_dispatches(root, counter):
    """Per kernel, the counter's value of every dispatch (rows of one
    dispatch summed)."""
    per = {}
    for r in _csv_rows(root, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        k = _short_kernel(r["Kernel_Name"])
        d = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per.setdefault(k, {})
        per[k][d] = per[k].get(d, 0.0) + float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in per.items()}


def _counter_per_dispatch(root, counter):
    """Counter value per dispatch, averaged over each kernel's dispatches."""
    return {k: (sum(v) / len(v), len(v)) for k, v in _counter_dispatches(root, counter).items()}


def _copy_kib(root, counter, copies=5):
    """The counter per 1 GiB copy: the mean of the `copies` largest copy
    dispatches (the probe's small read-back uses the copy kernel too)."""
    vals = [v for k, vs in _counter_dispatches(root, counter).items()
            if "copyBuffer" in k for v in vs]
    vals = sorted(vals)[-copies:]
    return sum(vals) / len(vals)


def _rocprof_run(name, outdir, tag, prof, bench_args, timeout, summarize_only):
    """One rocprofv3 run of this script with `bench_args`; raw output in
    outdir/<name>_<tag>.  Returns that folder, the program's last JSON line
    and the command as the summary shows it."""
    me = os.path.abspath(__file__)
    d = os.path.join(os.path.abspath(outdir), f"{name}_{tag}")  # rocprofv3 runs in /tmp
    cmd = ["rocprofv3", *prof, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, me, *bench_args]
    stdout_file = os.path.join(d, "stdout.txt")
    if not summarize_only:
        print(f"[profile] {name}: {' '.join(cmd)}", file=sys.stderr, flush=True)
        # stdout (the JSON line) is captured; the program's progress lines
        # and the profiler's messages go on to this process's stderr
        r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                           stdout=subprocess.PIPE, text=True, timeout=timeout)
        last = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 and not last:
            sys.stderr.write(r.stdout[-2000:])
            raise RuntimeError(f"rocprofv3 {tag} failed ({r.returncode})")
        os.makedirs(d, exist_ok=True)
        with open(stdout_file, "w") as f:
            f.write(r.stdout)
    with open(stdout_file) as f:
        last = [l for l in f.read().splitlines() if l.startswith("{")]
    return d, json.loads(last[-1]) if last else None, " ".join(
        ["rocprofv3", *prof, "--output-format csv", "-d", f"<dir>/{name}_{tag}", "-o run --",
         "python3 bench.py", *bench_args])


def _workload_launch(out):
    """Sites per dispatch and the HIP-event average launch time that the
    profiled workload reports."""
    if "by_kernel" in out:
        v = next(iter(out["by_kernel"].values()))
        return v["sites_per_launch"], v["avg_launch_us"]
    return out["sites_per_launch"], out.get("avg_launch_us")


# Instruction and LDS counters of --rocprof-sq (eight SQ counters, one pass).
SQ_COUNTERS = ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
               "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES"]
# VALU issue capacity: 256 CUs x 4 SIMDs at 2.4 GHz, a wave64 VALU
# instruction holding a 16-lane SIMD for 4 cycles.
SIMD_WAVE_INSTS_PER_US = 256 * 4 * 2400 / 4


def rocprof_sq(name, outdir, summarize_only=False):
    """Runs rocprofv3 on one profiling workload with the SQ instruction and
    LDS counters (one pass) and a kernel trace (for the dispatch time), and
    writes profiles/<kernel>_<workload>_sq.md: per dispatch of the
    workload's kernel, wave-level VALU / SALU / LDS instruction counts, VALU
    lane-instructions per site, the share of the chip's VALU issue cycles
    the kernel uses and LDS bank conflicts."""
    args, kernel, _ = PROFILE_WORKLOADS[name]
    os.makedirs(outdir, exist_ok=True)
    kt_dir, out, kt_cmd = _rocprof_run(name, outdir, "sqkt", ["--kernel-trace", "--stats"], args,
                                       600, summarize_only)
    sq_dir, _, sq_cmd = _rocprof_run(name, outdir, "sq", ["--pmc", *SQ_COUNTERS], args, 600,
                                     summarize_only)
    stats = _kernel_stats(kt_dir)
    if kernel not in stats:
        raise RuntimeError(f"{kernel} not in the kernel trace: {sorted(stats)}")
    sites, ev_us = _workload_launch(out)
    avg_us = stats[kernel]["avg_us"]
    per = {c: _counter_per_dispatch(sq_dir, c).get(kernel, (0.0, 0)) for c in SQ_COUNTERS}
    valu, nd = per["SQ_INSTS_VALU"]
    lds = per["SQ_INSTS_LDS"][0]
    lines = [f"# {kernel} -- {name}, instruction and LDS counters", "",
             f"Workload: `python3 bench.py {' '.join(args)}`: {PROFILE_TEXT[name]}.", "",
             f"Commands (made by `python3 bench.py --rocprof-sq {name} <dir>`):", "", "```",
             "cd /tmp && export TMPDIR=/tmp", kt_cmd, sq_cmd, "```", "",
             f"Per dispatch of `{kernel}` (means over {nd} dispatches; kernel trace of the same "
             f"workload: {stats[kernel]['calls']} dispatches, average {avg_us:.2f} us"
             + (f", HIP events {ev_us:.2f} us" if ev_us else "") + "):", "",
             "| Counter | per dispatch |", "|---|---|"]
    for c in SQ_COUNTERS:
        lines.append(f"| {c} | {per[c][0] / 1e6:.3f} M |")
    lines += ["",
              f"* VALU lane-instructions per site per dispatch: {64 * valu / sites:.1f} "
              f"({sites} sites per dispatch).",
              f"* VALU issue: {100 * valu / (SIMD_WAVE_INSTS_PER_US * avg_us):.1f} % of the "
              "chip's SIMD cycles over the dispatch (256 CUs x 4 SIMDs x 2.4 GHz, 4 cycles "
              "per wave64 VALU instruction).",
              f"* LDS instructions: {lds / 1e6:.3f} M per dispatch, "
              f"{per['SQ_LDS_BANK_CONFLICT'][0] / 1e6:.3f} M bank-conflict cycles."]
    md = f"{kernel.replace('<', '_').replace('>', '').replace(', ', '_')}_{name}_sq.md"
    with open(os.path.join(ROOT, "profiles", md), "w") as f:
        f.write("\n".join(lines) + "\n")
    # wave-level VALU instructions per site per dispatch, which the headline
    # line's roofline.valu scales by its own sites and launch time
    jpath = os.path.join(ROOT, "profiles", "kernels_sq.json")
    table = {}
    if os.path.exists(jpath):
        with open(jpath) as f:
            table = json.load(f)
    profiled = {k for _, k, _ in PROFILE_WORKLOADS.values()}
    table = {k: v for k, v in table.items() if k in profiled}
    table[kernel] = {"workload": name, "valu_wave_insts_per_site": round(valu / sites, 4),
                     "sites_per_dispatch": sites, "avg_launch_us": round(avg_us, 2),
                     "valu_issue_frac": round(valu / (SIMD_WAVE_INSTS_PER_US * avg_us), 4),
                     "profile": md}
    with open(jpath, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"rocprof_sq": name, "kernel": kernel, "avg_us": avg_us,
                      "valu_per_dispatch": valu, "summary": md}), flush=True)


# The driver's own invocation of this benchmark, traced by --rocprof-bench.
BENCH_DEFAULT_ARGS = ["--gpus", "1", "--steps", "20", "--warmup", "5"]


def rocprof_bench(outdir, summarize_only=False):
    """Runs the default benchmark command under rocprofv3 --kernel-trace
    --stats and writes profiles/bench_default_kernel_trace.md: the kernel
    statistics of the whole run, and the headline kernel's dispatches inside
    the headline section (every dispatch before the plain path's first
    k_lozenge_step), whose average duration is the one the JSON line's
    roofline uses (roofline.avg_launch_us, HIP events on the same
    launches).  Raw output stays in `outdir` (the dispatch trace of the
    full run is hundreds of MB)."""
    import csv
    name = "bench_default"
    d, out, cmd = _rocprof_run(name, outdir, "kt", ["--kernel-trace", "--stats"],
                               BENCH_DEFAULT_ARGS, 1100, summarize_only)
    roof = out["roofline"]
    kernel = roof["kernel"]
    stats = _kernel_stats(d)
    trace = []
    import glob
    for path in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                trace.append((int(r["Dispatch_Id"]), _short_kernel(r["Kernel_Name"]),
                              int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    trace.sort()
    head = []
    for _, k, t0, t1 in trace:
        if k.startswith("k_lozenge_step"):
            break
        if k == kernel:
            head.append((t0, t1))
    avg_head = sum(t1 - t0 for t0, t1 in head) / len(head) / 1e3
    # the batch's units run in parts on concurrent streams: the time the
    # kernel occupies the chip (union of the dispatch intervals) per launch
    # window, which is what the JSON line's HIP events time
    busy, end = 0, None
    for t0, t1 in sorted(head):
        if end is None or t0 > end:
            busy += t1 - t0
            end = t1
        elif t1 > end:
            busy += t1 - end
            end = t1
    busy_per_window = busy / roof["launches"] / 1e3
    lines = [f"# {kernel} -- the default benchmark command, kernel trace", "",
             f"Workload: `python3 bench.py {' '.join(BENCH_DEFAULT_ARGS)}` (the whole default "
             f"run: headline batch `{out['config']['workload']}`, then the plain paths, C1, "
             "the plain sweep, the batch walks, the weighted / bibone section, the served "
             "stream and the CPU baseline).", "",
             "Command (made by `python3 bench.py --rocprof-bench <dir>`):", "", "```",
             "cd /tmp && export TMPDIR=/tmp", cmd, "```", "",
             "Kernel statistics of the whole run (`kernel_stats.csv`, top kernels by time):", "",
             "| Kernel | Calls | Total ms | Avg us | Min us | Max us | % |",
             "|---|---|---|---|---|---|---|"]
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"])[:10]:
        lines.append(f"| `{k}` | {v['calls']} | {v['total_ms']:.1f} | {v['avg_us']:.2f} | "
                     f"{v['min_us']:.2f} | {v['max_us']:.2f} | {v['pct']:.2f} |")
    lines += ["",
              f"Headline section (dispatches before the plain path's first `k_lozenge_step`): "
              f"{len(head)} dispatches of `{kernel}` for {roof['launches']} launch windows "
              f"(the batch's samples in parts on concurrent streams, one dispatch per part "
              f"and window), average dispatch {avg_head:.2f} us; the kernel occupies the chip "
              f"(union of the dispatch intervals) {busy / 1e6:.1f} ms = "
              f"**{busy_per_window:.2f} us per launch window** by the trace, against the same "
              f"run's JSON line: avg_launch_us {roof['avg_launch_us']} per window (HIP events "
              f"around the windows), value {out['value']} {out['unit']}, roofline frac "
              f"{roof['frac']}.",
              f"Over the whole run the tracer counts {stats.get(kernel, {}).get('calls', 0)} "
              f"dispatches of `{kernel}`, average "
              f"{stats.get(kernel, {}).get('avg_us', 0.0):.2f} us (the served stream launches "
              "it too, on smaller hexagons)."]
    with open(os.path.join(ROOT, "profiles", "bench_default_kernel_trace.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps({"rocprof_bench": kernel, "headline_dispatches": len(head),
                      "trace_avg_dispatch_us": avg_head, "trace_us_per_window": busy_per_window,
                      "events_avg_us": roof["avg_launch_us"]}), flush=True)


def rocprof_profile(name, outdir, summarize_only=False):
    """Runs rocprofv3 on one profiling workload of this benchmark (kernel
    trace; FETCH_SIZE and WRITE_SIZE, each in a run of its own; the same two
    counters on a plain 1 GiB device copy, whose HBM bytes are known, to
    scale them) and writes profiles/<kernel>_<workload>.md and that kernel's
    entry of profiles/kernels_hbm.json (HBM bytes per site per dispatch,
    which the headline line's roofline.traffic uses).  Raw rocprofv3 output
    stays in `outdir`."""
    args, kernel, alg = PROFILE_WORKLOADS[name]
    os.makedirs(outdir, exist_ok=True)

    def run(tag, prof, bench_args, timeout):
        return _rocprof_run(name, outdir, tag, prof, bench_args, timeout, summarize_only)

    kt_dir, out, kt_cmd = run("kt", ["--kernel-trace", "--stats"], args, 600)
    f_dir, _, f_cmd = run("fetch", ["--pmc", "FETCH_SIZE"], args, 600)
    w_dir, _, w_cmd = run("write", ["--pmc", "WRITE_SIZE"], args, 600)
    cf_dir, _, cf_cmd = run("copy_fetch", ["--pmc", "FETCH_SIZE"], ["--copy-probe", "1024"], 300)
    cw_dir, _, cw_cmd = run("copy_write", ["--pmc", "WRITE_SIZE"], ["--copy-probe", "1024"], 300)
    stats = _kernel_stats(kt_dir)
    fetch, write = _counter_per_dispatch(f_dir, "FETCH_SIZE"), _counter_per_dispatch(w_dir,
                                                                                       "WRITE_SIZE")
    copy = "__amd_rocclr_copyBuffer"
    moved_kib = 1024 * 1024  # 1 GiB read and 1 GiB written per copy
    f_factor = moved_kib / _copy_kib(cf_dir, "FETCH_SIZE")
    w_factor = moved_kib / _copy_kib(cw_dir, "WRITE_SIZE")
    if kernel not in stats:
        raise RuntimeError(f"{kernel} not in the kernel trace: {sorted(stats)}")
    # sites per dispatch of the kernel, from the workload's own report
    sites, ev_us = _workload_launch(out)
    st = stats[kernel]
    fk, fn = fetch[kernel]
    wk, wn = write[kernel]
    hbm = (fk * f_factor + wk * w_factor) * 1024
    alg_b = alg * sites
    lines = [f"# {kernel} -- {name}", "",
             f"Workload: `python3 bench.py {' '.join(args)}`: {PROFILE_TEXT[name]}.",
             "", "Commands (made by `python3 bench.py --rocprof-profile " + name +
             " <dir>`; each counter in a run of its own):", "", "```",
             "cd /tmp && export TMPDIR=/tmp", kt_cmd, f_cmd, w_cmd, cf_cmd, cw_cmd, "```", "",
             "Kernel trace (`kernel_stats.csv`, top kernels):", "",
             "| Kernel | Calls | Total ms | Avg us | Min us | Max us | % |",
             "|---|---|---|---|---|---|---|"]
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"])[:6]:
        lines.append(f"| `{k}` | {v['calls']} | {v['total_ms']:.1f} | {v['avg_us']:.2f} | "
                     f"{v['min_us']:.2f} | {v['max_us']:.2f} | {v['pct']:.2f} |")
    lines += ["", f"`{kernel}`: {st['calls']} dispatches, average {st['avg_us']:.2f} us"
              + (f" (HIP events in the same program: {ev_us:.2f} us per launch)." if ev_us else "."),
              f"Algorithmic bytes: {alg} B per site x {sites} sites per dispatch = "
              f"{alg_b / 1e6:.2f} MB, {alg_b / st['avg_us'] / 1e3:.0f} GB/s = "
              f"{100 * alg_b / st['avg_us'] / 1e3 / HBM_PEAK_GBS:.1f} % of 8 TB/s.", "",
              f"HBM counters per dispatch, scaled by a plain 1 GiB device copy "
              f"(`{copy}`: FETCH_SIZE factor {f_factor:.4f}, WRITE_SIZE factor {w_factor:.4f}):", "",
              "| Dispatches (fetch / write) | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled HBM bytes | "
              "bytes per site |", "|---|---|---|---|---|",
              f"| {fn} / {wn} | {fk:.1f} | {wk:.1f} | {hbm / 1e6:.2f} MB | {hbm / sites:.1f} |", "",
              f"Measured HBM traffic {hbm / st['avg_us'] / 1e3:.0f} GB/s "
              f"({100 * hbm / st['avg_us'] / 1e3 / HBM_PEAK_GBS:.1f} % of peak)."]
    md = f"{kernel.replace('<', '_').replace('>', '').replace(', ', '_')}_{name}.md"
    with open(os.path.join(ROOT, "profiles", md), "w") as f:
        f.write("\n".join(lines) + "\n")
    jpath = os.path.join(ROOT, "profiles", "kernels_hbm.json")
    table = json.load(open(jpath)) if os.path.exists(jpath) else {}
    # entries of kernels this mode profiles only
    profiled = {k for _, k, _ in PROFILE_WORKLOADS.values()}
    table = {k: v for k, v in table.items() if k in profiled}
    table[kernel] = {"workload": name, "fetch_kib_per_dispatch": round(fk, 2),
                     "write_kib_per_dispatch": round(wk, 2), "fetch_factor": round(f_factor, 5),
                     "write_factor": round(w_factor, 5), "sites_per_dispatch": sites,
                     "hbm_bytes_per_site": round(hbm / sites, 3),
                     "avg_launch_us": round(st["avg_us"], 2), "profile": md}
    with open(jpath, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"rocprof_profile": name, "kernel": kernel, "avg_us": st["avg_us"],
                      "hbm_bytes_per_site": hbm / sites, "summary": md}), flush=True)


def profile_rounds(ctx, rt, tmax, tmin, masters, t0, rounds):
    """The benchmark's batch, stopped after `rounds` doubling rounds: every
    launch has the timed run's shape, so per-launch kernel time matches."""
    w = time.time()
    try:
        ctx.lozenge_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, t0, rounds)
    except rt.RTError:
        pass  # no coalescence within the cap is expected here
    tm = ctx.last_timing()
    kl = max(tm["kernel_launches"], 1)
    by_kernel = {}
    for k, name in enumerate(KERNEL_BY_SHAPE):
        n = tm["shape_launches"][k]
        if name and n:
            by_kernel[name] = {"launches": n, "avg_launch_us": round(tm["shape_ms"][k] * 1e3 / n, 2),
                               "sites_per_launch": int(tm["shape_sites"][k] / n)}
    print(json.dumps({"profile_rounds": rounds, "samples": len(masters), "wall_s": time.time() - w,
                      "kernel_ms": tm["kernel_ms"], "kernel_launches": tm["kernel_launches"],
                      "avg_launch_us": round(tm["kernel_ms"] * 1e3 / kl, 2),
                      "sites_per_launch": int(tm["kernel_sites"] / kl),
                      "steps": tm["steps"], "by_kernel": by_kernel}), flush=True)


C1_RULES = (("toggle08", 0), ("heatbath", 1))


def c1_gpu(rt, steps=10 ** 4):
    """The reference's CPU-runnable configuration on the device: Walk of
    MinTiling(AztecDiamond(100)) (N = 206), 10^4 steps, seed 1, with the 0.8
    toggle (dominokernel.cl) and the heat-bath rule (dominokernelCFTP.cl),
    in the reference's launch shape (one launch per colour step) and on the
    sweep-block kernel.  Footprint 0.7 MB: cache-resident, launch-bound.
    Then the height function and edge counts of the result
    (DominoTiler.cpp:555-593, DominoTiler.h:28-32).  Returns the results too,
    so the CPU leg can check its walks against them."""
    d = rt.DominoTiler.AztecDiamond(100)
    t = rt.DominoTiler.MinTiling(d)
    N = rt.side(t)
    ctx = rt.Context(0)
    out, results = {"workload": f"domino Aztec diamond order 100 (N={N}), Walk 10^4 steps, "
                                "seed 1, from the minimal tiling; cache-resident"}, {}
    sites = (N - 2) * (N // 2 - 2)
    try:
        ctx.set_tinymt_params(rt.synthetic_params(N * N // 2, None))
        for name, rule in C1_RULES:
            rec = {}
            for eng_name, eng in (("one_launch_per_step", rt.ENGINE_STEP), ("sweep_block", 0)):
                ctx.set_engine(eng)
                ctx.domino_walk(t, 200, 1, rule)
                w = time.time()
                res = ctx.domino_walk(t, steps, 1, rule)
                wall = time.time() - w
                tm = ctx.last_timing()
                rec[eng_name] = {"wall_s": round(wall, 4),
                                 "device_s": round(tm["device_ms"] / 1e3, 4),
                                 "flip_attempts_per_s": round(steps * sites / wall, 1),
                                 "launches": tm["launches"]}
                if name in results:
                    assert np.array_equal(results[name], res)
                results[name] = res
            hf = rt.DominoTiler.TilingToHeightfunc(results[name], d)
            r = results[name]
            rec["edge_counts"] = [int(((r >> b) & 1).sum()) for b in range(4)]
            rec["height_range"] = [int(hf[hf != np.iinfo(np.int32).max].min()),
                                   int(hf[hf != np.iinfo(np.int32).max].max())]
            out[name] = rec
        ctx.set_engine(0)
    finally:
        ctx.close()
    return out, results


def cpu_c1(threads, gpu_results=None, steps=10 ** 4):
    """The reference's CPU-runnable case (Aztec diamond of order 100, seed
    1, all 10^4 steps) on the oracle, with 1 thread and with `threads`
    threads, for both rules; each walk must equal the device's."""
    import randomtilingspublic_amd as rt
    from tests import oracle_lib as o
    d = rt.DominoTiler.AztecDiamond(100)
    t = rt.DominoTiler.MinTiling(d)
    N = rt.side(t)
    p = rt.synthetic_params(N * N // 2, None)
    out = {"workload": f"domino Aztec diamond order 100, seed 1, {steps} steps"}
    sites = (N - 2) * (N // 2 - 2)
    for name, rule in C1_RULES:
        rec = {}
        for th in (1, threads):
            o.set_threads(th)
            t0 = time.time()
            res = o.domino_walk(t, steps, 1, p, rule)
            dt = time.time() - t0
            rec[f"threads_{th}"] = {"flip_attempts_per_s": round(steps * sites / dt, 1),
                                    "s": round(dt, 3)}
            if gpu_results is not None:
                assert np.array_equal(res, gpu_results[name]), name
        out[name] = rec
    o.set_threads(1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32,
                    help="accepted for the harness; the timed batch is always master seeds "
                         "1..32 (see --samples)")
    ap.add_argument("--samples", type=int, default=32,
                    help="profiling only: master seeds 1..K in the batch (32 = the full "
                         "configuration, the only one reported as the headline)")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=300)
    ap.add_argument("--t0", type=int, default=1024)
    ap.add_argument("--max-rounds", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batch-walks", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-cftp", type=int, default=8,
                    help="two-periodic Aztec CFTP for master seeds 1..K (0: skip)")
    ap.add_argument("--profile-rounds", type=int, default=0,
                    help="profiling run: the same batch capped at R doubling rounds (no sample "
                         "needs to finish); prints the sweep kernel's timing only")
    ap.add_argument("--profile-section", choices=["sweep", "walks", "c3", "c3scan", "bibone",
                                                  "served"],
                    default=None,
                    help="profiling run: only one section (plain Aztec sweep, 256-chain batch "
                         "walks, weighted walk / bibone, served stream), printed as one JSON "
                         "line")
    ap.add_argument("--serve-dom", type=int, nargs=2, default=[64, 256], metavar=("LO", "HI"),
                    help="served stream: domino Aztec orders")
    ap.add_argument("--serve-loz", type=int, nargs=2, default=[64, 160], metavar=("LO", "HI"),
                    help="served stream: lozenge hexagon sides")
    ap.add_argument("--serve-rt", type=int, nargs=2, default=[32, 256], metavar=("LO", "HI"),
                    help="served stream: rectangle-triangle hexagon sides of the walk requests "
                         "(0 0: none)")
    ap.add_argument("--serve-rt-steps", type=int, default=10 ** 4,
                    help="served stream: steps of each rectangle-triangle walk request")
    ap.add_argument("--serve-seconds", type=float, default=40.0)
    ap.add_argument("--serve-plain", type=int, default=6,
                    help="served stream: requests timed on the plain paths")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="served stream: arrival rate as a fraction of the burst capacity")
    ap.add_argument("--serve-burst", type=int, default=32,
                    help="served stream: requests of the burst that measures capacity")
    ap.add_argument("--bibone-steps", type=int, default=10 ** 5,
                    help="bibone profiling section: walk steps")
    ap.add_argument("--xcd-order", type=int, default=None, choices=[0, 1, 2],
                    help="rt_set_xcd_order for the headline batch and the batch walks "
                         "(default: the library's)")
    ap.add_argument("--halo", type=int, default=0,
                    help="sweep-block halo depth (rt_set_halo; 0 = the library's default; "
                         "results never depend on it)")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams the headline batch and the batch walks split their units "
                         "over (rt_set_streams; 0 = the library's default, 2); the per-dispatch "
                         "profiles use 1")
    ap.add_argument("--tile-shape", type=int, default=0,
                    help="force a sweep-block tile shape (rt_set_tile_shape; 0 = autotuned), "
                         "for profiling one kernel")
    ap.add_argument("--rocprof-profile", nargs=2, metavar=("WORKLOAD", "OUTDIR"), default=None,
                    help="run rocprofv3 (kernel trace, FETCH_SIZE, WRITE_SIZE, copy calibration) "
                         f"on one of {sorted(PROFILE_WORKLOADS)} and write its summary under "
                         "profiles/; raw output in OUTDIR")
    ap.add_argument("--rocprof-sq", nargs=2, metavar=("WORKLOAD", "OUTDIR"), default=None,
                    help="run rocprofv3 with instruction and LDS counters (SQ_*) on one "
                         "profiling workload and write profiles/<kernel>_<workload>_sq.md")
    ap.add_argument("--rocprof-bench", metavar="OUTDIR", default=None,
                    help="run the default benchmark command under rocprofv3 --kernel-trace "
                         "--stats and write profiles/bench_default_kernel_trace.md")
    ap.add_argument("--summarize-only", action="store_true",
                    help="with --rocprof-profile / --rocprof-sq: only re-read the raw output "
                         "in OUTDIR")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB: time a plain device copy of that size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only (independent samples "
                                   "shard across GPUs without communication; not measured here)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    global XCD_ORDER, STREAMS
    XCD_ORDER = args.xcd_order
    STREAMS = args.streams
    if args.rocprof_profile:
        rocprof_profile(*args.rocprof_profile, summarize_only=args.summarize_only)
        return
    if args.rocprof_sq:
        rocprof_sq(*args.rocprof_sq, summarize_only=args.summarize_only)
        return
    if args.rocprof_bench:
        rocprof_bench(args.rocprof_bench, summarize_only=args.summarize_only)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import randomtilingspublic_amd as rt

    def serve():
        return served_stream(rt, seconds=args.serve_seconds, dom=tuple(args.serve_dom),
                             loz=tuple(args.serve_loz), plain_n=args.serve_plain,
                             burst_n=args.serve_burst, load=args.serve_load,
                             rtri=tuple(args.serve_rt) if args.serve_rt[1] > 0 else None,
                             rtri_steps=args.serve_rt_steps)

    if args.profile_section:
        fn = {"sweep": lambda: plain_sweep(rt), "walks": lambda: batched_walks(rt),
              "c3": lambda: weighted_and_bibone(rt), "bibone": lambda: bibone_section(rt, args.tile_shape, args.bibone_steps),
              "c3scan": lambda: two_periodic_scan(rt),
              "served": serve}[args.profile_section]
        print(json.dumps({"profile_section": args.profile_section, **fn()}), flush=True)
        return

    set_halo_kernels(args.halo)
    side = args.side
    d = rt.LozengeTiler.AlmostHexDomain(side, 0)
    tmin, tmax = rt.LozengeTiler.MinTiling(d), rt.LozengeTiler.MaxTiling(d)
    N = rt.side(tmin)
    P = N * (N // 3)
    params = rt.synthetic_params(N * N, None)
    ctx = rt.Context(0)
    ctx.set_tinymt_params(params)
    ctx.set_tile_shape(args.tile_shape)
    ctx.set_halo(args.halo)
    if STREAMS:
        ctx.set_streams(STREAMS)
    if XCD_ORDER is not None:
        ctx.set_xcd_order(XCD_ORDER)
    print(f"[bench] device {ctx.device_name} hexagon {side} N={N} sites/step={P}", file=sys.stderr,
          flush=True)
    for w in range(args.warmup):
        ctx.lozenge_walk(tmin, 2048, 1000 + w)
    masters = list(range(1, args.samples + 1))
    if args.profile_rounds:
        profile_rounds(ctx, rt, tmax, tmin, masters, args.t0, args.profile_rounds)
        return
    ctx.set_round_callback(lambda r, n, fin, tot: print(
        f"[bench] round {r} ({n} steps): {fin} of {tot} samples coalesced", file=sys.stderr,
        flush=True))
    t0 = time.time()
    outs, stats = ctx.lozenge_cftp_batch(tmax, tmin, masters, rt.SCHED_DOUBLING, args.t0,
                                         args.max_rounds)
    wall = time.time() - t0
    ctx.set_round_callback(None)
    tm = ctx.last_timing()
    dev_s = tm["device_ms"] / 1e3
    draws = tm["attempts"]                  # one TinyMT draw per site per step per sample
    attempts = 2 * draws                    # two chains (top, bottom) per draw
    samples_s = len(masters) / wall
    print(f"[bench] batched CFTP {len(masters)} samples wall {wall:.2f}s device {dev_s:.2f}s "
          f"rounds {[s['rounds'] for s in stats]}", file=sys.stderr, flush=True)
    # exactness guard: every sample is a tiling of the hexagon
    dom = rt.LozengeTiler.TilingToDomain(tmin)
    for k in range(len(masters)):
        assert np.array_equal(rt.LozengeTiler.TilingToDomain(outs[k]), dom)

    # Plain paths: one sample per call.  "plain" is the reference's launch
    # shape (one launch per colour step, ENGINE_STEP); "plain_fused" is the
    # same call on this engine's sweep-block kernel.  Both must return the
    # batched run's sample bit for bit.
    plain = plain_fused = None
    if not args.no_plain:
        for engine in (rt.ENGINE_STEP, 0):
            ctx.set_engine(engine)
            t1 = time.time()
            o1, s1 = ctx.lozenge_cftp(tmax, tmin, masters[0], rt.SCHED_DOUBLING, args.t0,
                                      args.max_rounds)
            pw = time.time() - t1
            assert np.array_equal(o1, outs[0])
            pt = ctx.last_timing()
            rec = {"samples_per_s": 1.0 / pw, "seed": masters[0], "wall_s": round(pw, 3),
                   "rounds": s1["rounds"],
                   "engine": "one launch per colour step" if engine else "sweep-block kernel"}
            if not engine:
                rec.update(tile_shape=pt["shape"], split_pairs=bool(pt["split"]))
            if engine:
                plain = rec
            else:
                plain_fused = rec
            print(f"[bench] plain CFTP ({rec['engine']}) seed {masters[0]} wall {pw:.2f}s",
                  file=sys.stderr, flush=True)
        ctx.set_engine(0)

    # Roofline of the sweep kernel with the most kernel time (the batch picks
    # a tile shape per round from its tile count): algorithmic bytes per
    # launch = for every sample's sites, TinyMT state (16 B) + parameters
    # (12 B) read, state written (16 B), and 3 colour planes x 2 chains read
    # and written (12 B).
    shape = int(np.argmax(tm["shape_ms"]))
    kname = KERNEL_BY_SHAPE[shape]
    klaunch = max(tm["shape_launches"][shape], 1)
    kms = tm["shape_ms"][shape]
    avg_launch_s = kms / 1e3 / klaunch
    sites_per_launch = tm["shape_sites"][shape] / klaunch
    alg_bytes = float(ALG_BYTES_BY_SHAPE[shape]) * sites_per_launch
    achieved_gbs = alg_bytes / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = None
    counters = os.path.join(ROOT, "profiles", "kernels_hbm.json")
    if os.path.exists(counters) and side == 300:
        # FETCH_SIZE / WRITE_SIZE per dispatch of this kernel on this workload,
        # scaled by a plain copy's readings, as bytes per site: times this
        # run's sites per launch.
        with open(counters) as f:
            per_kernel = json.load(f)
        if kname in per_kernel:
            traffic = int(per_kernel[kname]["hbm_bytes_per_site"] * sites_per_launch)
    # The fused kernel runs about 14 colour steps per launch from LDS, so
    # HBM does not bind it: beside the HBM figure, its VALU issue -- wave
    # instructions per site per launch from the committed SQ-counter profile
    # of the same kernel and workload, times this run's sites per launch,
    # over this run's launch time and the chip's issue capacity.
    valu = None
    sqfile = os.path.join(ROOT, "profiles", "kernels_sq.json")
    if os.path.exists(sqfile) and side == 300:
        with open(sqfile) as f:
            per_sq = json.load(f)
        if kname in per_sq and avg_launch_s > 0:
            ins = per_sq[kname]["valu_wave_insts_per_site"] * sites_per_launch
            valu = {"issue_frac": round(ins / (SIMD_WAVE_INSTS_PER_US * avg_launch_s * 1e6), 4),
                    "wave_insts_per_launch": int(ins),
                    "peak_wave_insts_per_us": SIMD_WAVE_INSTS_PER_US,
                    "source": "profiles/kernels_sq.json (" + per_sq[kname]["profile"] + ")"}
    roof = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "alg_bytes_per_site": ALG_BYTES_BY_SHAPE[shape],
            "kernel_ms_by_shape": [round(x, 1) for x in tm["shape_ms"]],
            "avg_launch_us": round(avg_launch_s * 1e6, 2),
            "launches": klaunch, "kernel_time_share": round(kms / max(tm["kernel_ms"], 1e-9), 3),
            "sites_per_launch": int(sites_per_launch), "alg_bytes_per_launch": int(alg_bytes),
            "flip_attempts_per_s_in_kernels": round(2 * draws / (tm["kernel_ms"] / 1e3), 1)
            if tm["kernel_ms"] else None,
            # colour steps per launch and site (TinyMT draws over launched sites)
            "steps_per_launch": round(draws / tm["kernel_sites"], 2)
            if tm["kernel_sites"] else None,
            "launch_note": "a launch is one window of every unfinished sample; the samples run "
                           "in two parts on concurrent streams (rt_set_streams), one dispatch "
                           "per part, so avg_launch_us is HIP-event time per window",
            "valu": valu}
    # the later sections make contexts of their own; this one's streams
    # would share the process's hardware queues with theirs
    ctx.close()
    sweep = None if args.no_sweep else plain_sweep(rt)
    if sweep:
        print(f"[bench] plain sweep {sweep}", file=sys.stderr, flush=True)
    walks = None if args.no_batch_walks else batched_walks(rt)
    if walks:
        print(f"[bench] batched walks {walks}", file=sys.stderr, flush=True)
    c3 = None if args.no_c3 else weighted_and_bibone(rt, args.c3_cftp)
    if c3:
        print(f"[bench] weighted / bibone {c3}", file=sys.stderr, flush=True)
    served = None
    if not args.no_serve:
        served = serve()
        print(f"[bench] served {served}", file=sys.stderr, flush=True)
    c1 = c1_results = None
    if not args.no_c1:
        c1, c1_results = c1_gpu(rt)
        print(f"[bench] C1 on the device {c1}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        # every CPU this process may run on (the box's share of the host)
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (
            os.cpu_count() or 1)
        cpu = cpu_baseline(tmax, tmin, params, attempts, len(masters), threads)
        cpu["nproc"] = os.cpu_count()
        cpu["cpu_model"] = cpu_model()
        cpu["c1"] = cpu_c1(threads, c1_results)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": "CFTP samples/s (flip attempts/s alongside), batched vs plain, 1 GPU",
        "value": round(samples_s, 4),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": len(masters),
        "steps_requested": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / len(masters), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (hexagon domain built in-process; synthetic TinyMT parameter table)",
        "config": {"workload": f"lozenge hexagon a=b=c={side}, CFTP from min/max, doubling from "
                               f"2^{int(np.log2(args.t0))}, master seeds 1..{len(masters)}, "
                               "one exact sample per seed, batched",
                   "N": N, "sites_per_step": P},
        "flip_attempts_per_s": round(attempts / wall, 1),
        "flip_attempts_per_s_device": round(attempts / dev_s, 1),
        "device_s": round(dev_s, 3),
        "rounds": [s["rounds"] for s in stats],
        "tile_shape_last_round": tm["shape"], "split_pairs_last_round": bool(tm["split"]),
        "plain": plain,
        "plain_fused": plain_fused,
        "batched_vs_plain": round(samples_s / plain["samples_per_s"], 2) if plain else None,
        "batched_vs_plain_fused":
            round(samples_s / plain_fused["samples_per_s"], 2) if plain_fused else None,
        "c1": c1,
        "plain_sweep": sweep,
        "batched_walks": walks,
        "weighted_and_bibone": c3,
        "served": served,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

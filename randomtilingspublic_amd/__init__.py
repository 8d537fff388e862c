"""MI355X random-tiling MCMC engine: Python view of librtiling.so.

The engine is C++/HIP (randomtilingspublic_amd/csrc) behind the C ABI in
include/rtiling.h; this module binds that ABI with ctypes so tests and the
benchmark can drive it.  It never computes a tiling itself and has no CPU
fallback: if librtiling.so is missing, every call raises.

Names follow the reference's per-model classes (DominoTiler, LozengeTiler,
RectTriangleTiler): Walk, LoadTinyMT, AztecDiamond, MinTiling, ...
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtiling.so")

RT_OK = 0
ERRORS = {
    -1: "EINVAL", -2: "EPARAMS", -3: "EIO", -4: "ENOCOUPLE", -5: "EDEVICE",
    -6: "ENOMEM", -7: "EUNTILEABLE", -8: "EADMIT", -9: "ECANCELLED", -10: "EAGAIN",
}

DOMINO_TOGGLE08, DOMINO_HEATBATH, DOMINO_METRO2P, DOMINO_HEATBATH2P = 0, 1, 2, 3
SCHED_REFERENCE, SCHED_DOUBLING = 0, 1
ENGINE_AUTO, ENGINE_STEP, ENGINE_FUSED, ENGINE_FUSED_SWAR = 0, 1, 2, 3


class RTError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class _Timing(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("launches", C.c_long),
                ("steps", C.c_long), ("attempts", C.c_long),
                ("kernel_ms", C.c_double), ("kernel_launches", C.c_long),
                ("kernel_sites", C.c_long), ("shape_ms", C.c_double * 8),
                ("shape_launches", C.c_long * 8), ("shape_sites", C.c_long * 8),
                ("shape", C.c_int), ("shape_source", C.c_int), ("split", C.c_int)]


class DominoRule(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_double), ("b", C.c_double)]


class CftpStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("total_steps", C.c_long),
                ("steps_run", C.c_long), ("device_ms", C.c_double)]


class Request(C.Structure):
    _fields_ = [("model", C.c_int), ("N", C.c_int), ("tmax", C.c_void_p), ("tmin", C.c_void_p),
                ("master_seed", C.c_uint32), ("sched", C.c_int), ("t0", C.c_long),
                ("max_rounds", C.c_int), ("rule", DominoRule), ("kind", C.c_int),
                ("steps", C.c_long)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_long), ("queued", C.c_long), ("ticks", C.c_long),
                ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double),
                ("span_s", C.c_double), ("samples_per_s", C.c_double), ("status", C.c_int)]


MODEL_DOMINO, MODEL_LOZENGE, MODEL_RECTTRI = 0, 1, 2
REQ_CFTP, REQ_WALK = 0, 1

_lib = None
_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_LONGP = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def lib():
    """Load librtiling.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run python -m randomtilingspublic_amd.build")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.rt_create.argtypes = [C.POINTER(vp), C.c_int]
    L.rt_destroy.argtypes = [vp]
    L.rt_destroy.restype = None
    L.rt_last_error.argtypes = [vp]
    L.rt_last_error.restype = C.c_char_p
    L.rt_device_name.argtypes = [vp]
    L.rt_device_name.restype = C.c_char_p
    L.rt_set_tinymt_params.argtypes = [vp, _U32P, C.c_size_t]
    L.rt_load_tinymt_file.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.rt_parse_tinymt_file.argtypes = [C.c_char_p, C.c_size_t, _U32P]
    L.rt_last_timing.argtypes = [vp, C.POINTER(_Timing)]
    L.rt_set_engine.argtypes = [vp, C.c_int]
    L.rt_domino_walk.argtypes = [vp, _I32P, C.c_int, C.c_long, C.c_long, C.POINTER(DominoRule)]
    L.rt_domino_walk_noalloc.argtypes = [vp, _I32P, C.c_int, C.c_long, C.c_long,
                                         C.POINTER(DominoRule)]
    L.rt_domino_walk_segments.argtypes = [vp, _I32P, C.c_int, _LONGP, _LONGP, C.c_int,
                                          C.POINTER(DominoRule)]
    L.rt_domino_cftp.argtypes = [vp, _I32P, _I32P, C.c_int, C.c_uint32, C.c_int, C.c_long,
                                 C.c_int, C.POINTER(DominoRule), _I32P, C.POINTER(CftpStats)]
    L.rt_lozenge_walk.argtypes = [vp, _I32P, C.c_int, C.c_long, C.c_long]
    L.rt_lozenge_walk_segments.argtypes = [vp, _I32P, C.c_int, _LONGP, _LONGP, C.c_int]
    L.rt_lozenge_cftp.argtypes = [vp, _I32P, _I32P, C.c_int, C.c_uint32, C.c_int, C.c_long,
                                  C.c_int, _I32P, C.POINTER(CftpStats)]
    L.rt_recttri_walk.argtypes = [vp, _I32P, C.c_int, C.c_long, C.c_long, C.c_void_p]
    L.rt_bibone_walk.argtypes = [vp, _I32P, C.c_int, C.c_long, C.c_long]
    L.rt_set_tile_shape.argtypes = [vp, C.c_int]
    L.rt_set_autotune.argtypes = [vp, C.c_int]
    L.rt_set_xcd_order.argtypes = [vp, C.c_int]
    L.rt_set_round_callback.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.rt_set_halo.argtypes = [vp, C.c_int]
    L.rt_set_streams.argtypes = [vp, C.c_int]
    L.rt_device_limits.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t)]
    L.rt_lozenge_cftp_batch.argtypes = [vp, _I32P, _I32P, C.c_int, _U32P, C.c_int, C.c_int,
                                        C.c_long, C.c_int, _I32P, C.c_void_p]
    L.rt_lozenge_cftp_batch_rounds.argtypes = [vp, _I32P, _I32P, C.c_int, _U32P, C.c_int,
                                               C.c_int, C.c_long, C.c_int, _I32P, _I32P,
                                               C.c_void_p]
    L.rt_domino_cftp_batch.argtypes = [vp, _I32P, _I32P, C.c_int, _U32P, C.c_int, C.c_int,
                                       C.c_long, C.c_int, C.POINTER(DominoRule), _I32P, C.c_void_p]
    L.rt_domino_cftp_batch_rounds.argtypes = [vp, _I32P, _I32P, C.c_int, _U32P, C.c_int,
                                              C.c_int, C.c_long, C.c_int, C.POINTER(DominoRule),
                                              _I32P, _I32P, C.c_void_p]
    L.rt_domino_walk_batch.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_long, _LONGP,
                                       C.POINTER(DominoRule), C.c_int]
    L.rt_domino_walk_batch_heights.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_long,
                                               _LONGP, C.POINTER(DominoRule), C.c_int, C.c_int,
                                               _LONGP, C.c_void_p]
    L.rt_domino_walk_batch_observe.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_long,
                                               _LONGP, C.POINTER(DominoRule), C.c_int, C.c_int,
                                               _LONGP]
    L.rt_server_create.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(vp)]
    L.rt_server_destroy.argtypes = [vp]
    L.rt_server_destroy.restype = None
    L.rt_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.rt_server_poll.argtypes = [vp, C.c_uint64, C.c_void_p, C.POINTER(CftpStats),
                                 C.POINTER(C.c_double)]
    L.rt_server_wait.argtypes = [vp, C.c_uint64, C.c_double]
    L.rt_server_cancel.argtypes = [vp, C.c_uint64]
    L.rt_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    # host helpers returning rt_vec*
    L.rt_vec_data.argtypes = [vp]
    L.rt_vec_data.restype = C.POINTER(C.c_int32)
    L.rt_vec_size.argtypes = [vp]
    L.rt_vec_size.restype = C.c_size_t
    L.rt_vec_free.argtypes = [vp]
    L.rt_vec_free.restype = None
    for name, args in [
        ("rt_domino_rectangle", [C.c_int, C.c_int]),
        ("rt_domino_aztec_diamond", [C.c_int]),
        ("rt_domino_aztec_rectangle", [C.c_int, C.c_void_p, C.c_int]),
        ("rt_domino_rect_union", [C.c_void_p, C.c_int]),
        ("rt_domino_heightfunc_to_tiling", [_I32P, C.c_size_t, _I32P, C.c_size_t]),
        ("rt_domino_tiling_to_heightfunc", [_I32P, C.c_size_t, _I32P, C.c_size_t]),
        ("rt_domino_tiling_to_domain", [_I32P, C.c_size_t]),
        ("rt_domino_domain_to_vertices", [_I32P, C.c_size_t]),
        ("rt_lozenge_almost_hex_domain", [C.c_int, C.c_int]),
        ("rt_lozenge_heightfunc_to_tiling", [_I32P, C.c_size_t, _I32P, C.c_size_t]),
        ("rt_lozenge_tiling_to_heightfunc", [_I32P, C.c_size_t, _I32P, C.c_size_t]),
        ("rt_lozenge_tiling_to_domain", [_I32P, C.c_size_t]),
        ("rt_lozenge_domain_to_vertices", [_I32P, C.c_size_t]),
        ("rt_recttri_max_hex", [C.c_int]),
        ("rt_recttri_min_hex", [C.c_int]),
        ("rt_recttri_slope_hex", [C.c_int]),
        ("rt_bibone_ice_cream_cone", [C.c_int, C.c_int]),
        ("rt_bibone_test_butterfly", [C.c_int]),
        ("rt_bibone_hexagon", [C.c_int, C.c_int, C.c_int]),
        ("rt_bibone_triangle_flips", [_I32P, C.c_size_t]),
        ("rt_bibone_butterfly_flips", [_I32P, C.c_size_t]),
        ("rt_bibone_tiling_to_vertices", [_I32P, C.c_size_t]),
        ("rt_bibone_tiling_to_domain", [_I32P, C.c_size_t]),
        ("rt_bibone_dimer_indicator", [_I32P, C.c_size_t, C.c_int]),
    ]:
        f = getattr(L, name)
        f.argtypes = args
        f.restype = vp
    L.rt_bibone_edge_form.argtypes = [_I32P, C.c_size_t, C.c_void_p]
    for name in ["rt_domino_domain_svg", "rt_domino_tiling_svg", "rt_domino_maya_svg",
                 "rt_lozenge_domain_svg", "rt_lozenge_tiling_svg", "rt_recttri_tiling_svg",
                 "rt_bibone_domain_svg", "rt_bibone_tiling_svg"]:
        getattr(L, name).argtypes = [_I32P, C.c_size_t, C.c_char_p]
    for name in ["rt_domino_dimer_svg", "rt_lozenge_dimer_svg", "rt_bibone_dimer_svg"]:
        getattr(L, name).argtypes = [_I32P, C.c_size_t, C.c_void_p, C.c_size_t, C.c_char_p]
    for name in ["rt_domino_max_tiling", "rt_domino_min_tiling",
                 "rt_lozenge_max_tiling", "rt_lozenge_min_tiling"]:
        getattr(L, name).argtypes = [_I32P, C.c_size_t, C.POINTER(vp)]
    _lib = L
    return L


def _take_vec(ptr):
    if not ptr:
        raise RTError(-1, "helper rejected its arguments")
    L = lib()
    n = L.rt_vec_size(ptr)
    out = np.ctypeslib.as_array(L.rt_vec_data(ptr), shape=(n,)).copy() if n else np.zeros(0, np.int32)
    L.rt_vec_free(ptr)
    return out.astype(np.int32)


def _svg(fn, a, path):
    a = _i32(a)
    rc = getattr(lib(), fn)(a, len(a), os.fsencode(path))
    if rc:
        raise RTError(rc, f"{fn}: cannot write {path}")


def _dimer_svg(fn, t, d, path):
    t = _i32(t)
    dd = None if d is None else _i32(d)
    rc = getattr(lib(), fn)(t, len(t), None if dd is None else dd.ctypes.data_as(C.c_void_p),
                            0 if dd is None else len(dd), os.fsencode(path))
    if rc:
        raise RTError(rc, f"{fn}: cannot write {path}")


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def side(a):
    n = int(round(np.sqrt(len(a))))
    if n * n != len(a):
        raise ValueError("array is not square")
    return n


_ROUND_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_long, C.c_int, C.c_int)


class Context:
    """One device, one stream, one TinyMT parameter table (rt_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.rt_create(C.byref(h), device)
        if rc != RT_OK:
            raise RTError(rc, "rt_create failed (no HIP device?)")
        self._h = h
        self._round_cb = None
        self._L = L

    def close(self):
        if self._h:
            self._L.rt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != RT_OK:
            raise RTError(rc, self._L.rt_last_error(self._h).decode())

    @property
    def device_name(self):
        return self._L.rt_device_name(self._h).decode()

    def set_engine(self, engine):
        self._chk(self._L.rt_set_engine(self._h, engine))

    def set_tile_shape(self, shape):
        """Sweep-block tile shape: 0 auto, 1 = 64 rows / 8 waves, 2 = 64 rows /
        16 waves, 3 = 32 rows / 16 waves; 4..6 = 1..3 with CFTP chain pairs
        split over two workgroups per tile; 7 = 48 rows / 16 waves at eight
        waves per SIMD."""
        self._chk(self._L.rt_set_tile_shape(self._h, shape))

    def device_limits(self):
        """(CUs, LDS bytes per CU, LDS bytes per workgroup)."""
        n, a, b = C.c_int(), C.c_size_t(), C.c_size_t()
        self._chk(self._L.rt_device_limits(self._h, C.byref(n), C.byref(a), C.byref(b)))
        return n.value, a.value, b.value

    def set_xcd_order(self, on):
        """Contiguous share of each launch's tiles per XCD (1, default) or
        the plain block order (0)."""
        self._chk(self._L.rt_set_xcd_order(self._h, on))

    def set_autotune(self, mode):
        """Tile-shape autotuner: 0 off, 1 on (default), 2 on and forget
        earlier choices."""
        self._chk(self._L.rt_set_autotune(self._h, mode))

    def set_halo(self, halo):
        """Sweep-block halo depth H (1..12; 0 = the default: 6, and 9 / 8
        for lozenge / domino chain-pair jobs on shape 7 whose 64-row tiles
        would fill over a third of the CUs): windows of at most H pulling
        steps; never changes results."""
        self._chk(self._L.rt_set_halo(self._h, halo))

    def set_streams(self, n):
        """Streams a batched CFTP / eager batched walk job is split over
        (default 2, 1..4); never changes results."""
        self._chk(self._L.rt_set_streams(self._h, n))

    def set_round_callback(self, fn):
        """fn(round, round_steps, finished, total) after each CFTP round's
        coalescence test (rt_set_round_callback); None turns it off."""
        if fn is None:
            self._round_cb = None
            self._chk(self._L.rt_set_round_callback(self._h, None, None))
            return
        cb = _ROUND_FN(lambda user, r, steps, fin, tot: fn(r, steps, fin, tot))
        self._round_cb = cb  # the C side holds only the pointer
        self._chk(self._L.rt_set_round_callback(self._h, C.cast(cb, C.c_void_p), None))

    def set_tinymt_params(self, params3):
        p = np.ascontiguousarray(np.asarray(params3, dtype=np.uint32).reshape(-1, 3).ravel())
        self._chk(self._L.rt_set_tinymt_params(self._h, p, len(p) // 3))

    def load_tinymt_file(self, path, count):
        self._chk(self._L.rt_load_tinymt_file(self._h, str(path).encode(), count))

    def last_timing(self):
        t = _Timing()
        self._chk(self._L.rt_last_timing(self._h, C.byref(t)))
        return {"device_ms": t.device_ms, "launches": t.launches, "steps": t.steps,
                "attempts": t.attempts, "kernel_ms": t.kernel_ms,
                "kernel_launches": t.kernel_launches, "kernel_sites": t.kernel_sites,
                "shape_ms": list(t.shape_ms), "shape_launches": list(t.shape_launches),
                "shape_sites": list(t.shape_sites), "shape": t.shape,
                "shape_source": t.shape_source, "split": t.split}

    # ---- domino ----
    def domino_walk(self, t, steps, seed, rule=DOMINO_TOGGLE08, a=1.0, b=1.0):
        t = _i32(t).copy()
        r = DominoRule(rule, a, b)
        self._chk(self._L.rt_domino_walk(self._h, t, side(t), steps, seed, C.byref(r)))
        return t

    def domino_walk_noalloc(self, t, steps, seed, rule=DOMINO_TOGGLE08, a=1.0, b=1.0):
        """domino_walk through the entry point that allocates nothing."""
        t = _i32(t).copy()
        r = DominoRule(rule, a, b)
        self._chk(self._L.rt_domino_walk_noalloc(self._h, t, side(t), steps, seed, C.byref(r)))
        return t

    def domino_walk_segments(self, t, steps, seeds, rule=DOMINO_TOGGLE08, a=1.0, b=1.0):
        t = _i32(t).copy()
        st = np.ascontiguousarray(steps, dtype=np.int64)
        se = np.ascontiguousarray(seeds, dtype=np.int64)
        r = DominoRule(rule, a, b)
        self._chk(self._L.rt_domino_walk_segments(self._h, t, side(t), st, se, len(st),
                                                  C.byref(r)))
        return t

    def domino_cftp(self, tmax, tmin, master_seed, sched=SCHED_REFERENCE, t0=128,
                    max_rounds=40, rule=DOMINO_HEATBATH, a=1.0, b=1.0):
        tmax, tmin = _i32(tmax), _i32(tmin)
        out = np.zeros_like(tmax)
        st = CftpStats()
        r = DominoRule(rule, a, b)
        self._chk(self._L.rt_domino_cftp(self._h, tmax, tmin, side(tmax), master_seed, sched,
                                         t0, max_rounds, C.byref(r), out, C.byref(st)))
        return out, {"rounds": st.rounds, "total_steps": st.total_steps,
                     "steps_run": st.steps_run, "device_ms": st.device_ms}

    # ---- lozenge ----
    def lozenge_walk(self, t, steps, seed):
        t = _i32(t).copy()
        self._chk(self._L.rt_lozenge_walk(self._h, t, side(t), steps, seed))
        return t

    def lozenge_walk_segments(self, t, steps, seeds):
        t = _i32(t).copy()
        st = np.ascontiguousarray(steps, dtype=np.int64)
        se = np.ascontiguousarray(seeds, dtype=np.int64)
        self._chk(self._L.rt_lozenge_walk_segments(self._h, t, side(t), st, se, len(st)))
        return t

    def lozenge_cftp(self, tmax, tmin, master_seed, sched=SCHED_REFERENCE, t0=128,
                     max_rounds=40):
        tmax, tmin = _i32(tmax), _i32(tmin)
        out = np.zeros_like(tmax)
        st = CftpStats()
        self._chk(self._L.rt_lozenge_cftp(self._h, tmax, tmin, side(tmax), master_seed, sched,
                                          t0, max_rounds, out, C.byref(st)))
        return out, {"rounds": st.rounds, "total_steps": st.total_steps,
                     "steps_run": st.steps_run, "device_ms": st.device_ms}

    # ---- batching ----
    def _cftp_batch(self, fn, tmax, tmin, masters, sched, t0, max_rounds, *rule, partial=False):
        tmax, tmin = _i32(tmax), _i32(tmin)
        masters = np.ascontiguousarray(masters, dtype=np.uint32)
        n = len(masters)
        out = np.zeros(n * len(tmax), dtype=np.int32)
        st = (CftpStats * n)()
        args = [self._h, tmax, tmin, side(tmax), masters, n, sched, t0, max_rounds]
        args += list(rule)
        args += [out, C.cast(st, C.c_void_p)]
        rc = fn(*args)
        stats = [{"rounds": s.rounds, "total_steps": s.total_steps, "steps_run": s.steps_run,
                  "device_ms": s.device_ms} for s in st]
        if not (partial and rc == -4):  # partial: keep the stats of an ENOCOUPLE batch
            self._chk(rc)
        return out.reshape(n, len(tmax)), stats

    def lozenge_cftp_batch(self, tmax, tmin, masters, sched=SCHED_DOUBLING, t0=1024,
                           max_rounds=40):
        return self._cftp_batch(self._L.rt_lozenge_cftp_batch, tmax, tmin, masters, sched, t0,
                                max_rounds)

    def _cftp_rounds(self, fn, tmax, tmin, masters, sched, t0, rounds, *rule):
        tmax, tmin = _i32(tmax), _i32(tmin)
        masters = np.ascontiguousarray(masters, dtype=np.uint32)
        n = len(masters)
        tops = np.zeros(n * len(tmax), dtype=np.int32)
        bots = np.zeros_like(tops)
        st = (CftpStats * n)()
        self._chk(fn(self._h, tmax, tmin, side(tmax), masters, n, sched, t0, rounds, *rule, tops,
                     bots, C.cast(st, C.c_void_p)))
        stats = [{"rounds": s.rounds, "total_steps": s.total_steps, "steps_run": s.steps_run,
                  "device_ms": s.device_ms} for s in st]
        return tops.reshape(n, len(tmax)), bots.reshape(n, len(tmax)), stats

    def lozenge_cftp_batch_rounds(self, tmax, tmin, masters, sched=SCHED_DOUBLING, t0=1024,
                                  rounds=1):
        """Both chains of every sample after exactly `rounds` CFTP rounds:
        (tops, bots, stats)."""
        return self._cftp_rounds(self._L.rt_lozenge_cftp_batch_rounds, tmax, tmin, masters,
                                 sched, t0, rounds)

    def domino_cftp_batch_rounds(self, tmax, tmin, masters, sched=SCHED_DOUBLING, t0=128,
                                 rounds=1, rule=DOMINO_HEATBATH, a=1.0, b=1.0):
        """Domino counterpart of lozenge_cftp_batch_rounds (heat-bath rules
        only): (tops, bots, stats)."""
        r = DominoRule(rule, a, b)
        return self._cftp_rounds(self._L.rt_domino_cftp_batch_rounds, tmax, tmin, masters,
                                 sched, t0, rounds, C.byref(r))

    def domino_cftp_batch(self, tmax, tmin, masters, sched=SCHED_DOUBLING, t0=128,
                          max_rounds=40, rule=DOMINO_HEATBATH, a=1.0, b=1.0, partial=False):
        """partial=True: when some samples do not coalesce within max_rounds,
        return anyway; those samples have stats total_steps == 0."""
        r = DominoRule(rule, a, b)
        return self._cftp_batch(self._L.rt_domino_cftp_batch, tmax, tmin, masters, sched, t0,
                                max_rounds, C.byref(r), partial=partial)

    def domino_walk_batch(self, tilings, steps, seeds, rule=DOMINO_TOGGLE08, a=1.0, b=1.0,
                          use_graph=False):
        ts = [_i32(t).copy() for t in tilings]
        Ns = np.array([side(t) for t in ts], dtype=np.int32)
        ptrs = (C.c_void_p * len(ts))(*[t.ctypes.data for t in ts])
        se = np.ascontiguousarray(seeds, dtype=np.int64)
        r = DominoRule(rule, a, b)
        self._chk(self._L.rt_domino_walk_batch(self._h, C.cast(ptrs, C.c_void_p),
                                               Ns.ctypes.data_as(C.c_void_p), len(ts), steps,
                                               se, C.byref(r), 1 if use_graph else 0))
        return ts

    def domino_walk_batch_observe(self, tilings, steps, seeds, rule=DOMINO_TOGGLE08, a=1.0,
                                  b=1.0, use_graph=False, download=True):
        """Batch walk plus device-side observables: returns (tilings or
        None, counts[nchains, 5]) with the vertices carrying edge bit 1, 2,
        4, 8 and the rotatable faces of each chain's final tiling."""
        ts = [_i32(t).copy() for t in tilings]
        Ns = np.array([side(t) for t in ts], dtype=np.int32)
        ptrs = (C.c_void_p * len(ts))(*[t.ctypes.data for t in ts])
        se = np.ascontiguousarray(seeds, dtype=np.int64)
        r = DominoRule(rule, a, b)
        counts = np.zeros((len(ts), 5), dtype=np.int64)
        self._chk(self._L.rt_domino_walk_batch_observe(
            self._h, C.cast(ptrs, C.c_void_p), Ns.ctypes.data_as(C.c_void_p), len(ts), steps, se,
            C.byref(r), 1 if use_graph else 0, 1 if download else 0, counts.reshape(-1)))
        return (ts if download else None), counts

    def domino_walk_batch_heights(self, tilings, steps, seeds, rule=DOMINO_TOGGLE08, a=1.0,
                                  b=1.0, use_graph=False, download=True):
        """Batch walk plus each chain's height function computed on the
        device (DominoTiler::TilingToHeightfunc of the final tiling on its
        own domain): returns (tilings or None, counts[nchains, 5],
        heights list of N x N int32 arrays flattened like a tiling)."""
        ts = [_i32(t).copy() for t in tilings]
        Ns = np.array([side(t) for t in ts], dtype=np.int32)
        ptrs = (C.c_void_p * len(ts))(*[t.ctypes.data for t in ts])
        hs = [np.zeros(len(t), dtype=np.int32) for t in ts]
        hptrs = (C.c_void_p * len(hs))(*[h.ctypes.data for h in hs])
        se = np.ascontiguousarray(seeds, dtype=np.int64)
        r = DominoRule(rule, a, b)
        counts = np.zeros((len(ts), 5), dtype=np.int64)
        self._chk(self._L.rt_domino_walk_batch_heights(
            self._h, C.cast(ptrs, C.c_void_p), Ns.ctypes.data_as(C.c_void_p), len(ts), steps, se,
            C.byref(r), 1 if use_graph else 0, 1 if download else 0, counts.reshape(-1),
            C.cast(hptrs, C.c_void_p)))
        return (ts if download else None), counts, hs

    # ---- rectangle-triangle ----
    def recttri_walk(self, t, steps, seed, weights16=None):
        t = _i32(t).copy()
        w = None
        if weights16 is not None:
            warr = np.ascontiguousarray(weights16, dtype=np.float32)
            assert warr.size == 16
            w = warr.ctypes.data_as(C.c_void_p)
        self._chk(self._L.rt_recttri_walk(self._h, t, side(t), steps, seed, w))
        return t

    def bibone_walk(self, t, steps, seed):
        """TriangleDimerTiler::Walk: returns the tiling after `steps` moves."""
        t = _i32(t).copy()
        self._chk(self._L.rt_bibone_walk(self._h, t, side(t), steps, seed))
        return t


class Server:
    """Continuous-batching CFTP server (rt_server) on a Context."""

    def __init__(self, ctx, hbm_budget, max_active=64):
        self._L = lib()
        self.ctx = ctx
        h = C.c_void_p()
        rc = self._L.rt_server_create(ctx._h, hbm_budget, max_active, C.byref(h))
        if rc != RT_OK:
            raise RTError(rc, "rt_server_create failed")
        self._h = h
        self._n = {}

    def close(self):
        if self._h:
            self._L.rt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, model, tmax, tmin, master_seed, sched=SCHED_DOUBLING, t0=1024,
               max_rounds=32, rule=DOMINO_HEATBATH, a=1.0, b=1.0):
        """A CFTP request; returns its id."""
        tmax, tmin = _i32(tmax), _i32(tmin)
        q = Request(model, side(tmax), tmax.ctypes.data, tmin.ctypes.data, master_seed, sched, t0,
                    max_rounds, DominoRule(rule, a, b), REQ_CFTP, 0)
        return self._submit(q, tmax)

    def submit_walk(self, model, start, steps, seed, rule=DOMINO_TOGGLE08, a=1.0, b=1.0):
        """A walk request: `steps` steps of the model's plain Walk from
        `start` with `seed` (rectangle-triangle: the reference kernel's
        weights); returns its id."""
        start = _i32(start)
        q = Request(model, side(start), start.ctypes.data, None, seed, SCHED_DOUBLING, 1024, 1,
                    DominoRule(rule, a, b), REQ_WALK, steps)
        return self._submit(q, start)

    def _submit(self, q, tmax):
        rid = C.c_uint64()
        rc = self._L.rt_server_submit(self._h, C.byref(q), C.byref(rid))
        if rc != RT_OK:
            raise RTError(rc, "submit rejected")
        self._n[rid.value] = len(tmax)
        return rid.value

    def poll(self, rid):
        """None while running; else (status, sample or None, stats, latency_s)."""
        out = np.zeros(self._n[rid], dtype=np.int32)
        st = CftpStats()
        lat = C.c_double()
        rc = self._L.rt_server_poll(self._h, rid, out.ctypes.data, C.byref(st), C.byref(lat))
        if rc == -10:
            return None
        del self._n[rid]
        stats = {"rounds": st.rounds, "total_steps": st.total_steps, "steps_run": st.steps_run}
        return rc, (out if rc == RT_OK else None), stats, lat.value

    def wait(self, rid, timeout_s=-1.0):
        return self._L.rt_server_wait(self._h, rid, timeout_s) == RT_OK

    def cancel(self, rid):
        rc = self._L.rt_server_cancel(self._h, rid)
        if rc != RT_OK:
            raise RTError(rc, "cancel")

    def stats(self):
        st = ServerStats()
        self._L.rt_server_get_stats(self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


def parse_tinymt_file(path, count):
    out = np.zeros(3 * count, dtype=np.uint32)
    rc = lib().rt_parse_tinymt_file(str(path).encode(), count, out)
    if rc != RT_OK:
        raise RTError(rc, f"cannot parse {path}")
    return out.reshape(count, 3)


# ---------------------------------------------------------------- statics
def _thurston(fn, d):
    d = _i32(d)
    h = C.c_void_p()
    rc = getattr(lib(), fn)(d, len(d), C.byref(h))
    if rc != RT_OK:
        raise RTError(rc, fn)
    return _take_vec(h.value)


class DominoTiler:
    """Static helpers of the reference's DominoTiler (DominoTiler.h:99-130)."""

    @staticmethod
    def Rectangle(M, N):
        return _take_vec(lib().rt_domino_rectangle(M, N))

    @staticmethod
    def AztecDiamond(n):
        return _take_vec(lib().rt_domino_aztec_diamond(n))

    @staticmethod
    def AztecRectangle(n, remove):
        r = np.ascontiguousarray(remove, dtype=np.int32)
        ptr = r.ctypes.data_as(C.c_void_p) if len(r) else None
        return _take_vec(lib().rt_domino_aztec_rectangle(n, ptr, len(r)))

    @staticmethod
    def RectangleUnion(rects):
        """Union of rectangles (i0, j0, h, w), framed like Rectangle, holes
        filled (no reference counterpart)."""
        r = np.ascontiguousarray(rects, dtype=np.int32).reshape(-1, 4)
        return _take_vec(lib().rt_domino_rect_union(r.ctypes.data_as(C.c_void_p), len(r)))

    @staticmethod
    def MaxTiling(d):
        return _thurston("rt_domino_max_tiling", d)

    @staticmethod
    def MinTiling(d):
        return _thurston("rt_domino_min_tiling", d)

    @staticmethod
    def HeightfuncToTiling(hf, d):
        hf, d = _i32(hf), _i32(d)
        return _take_vec(lib().rt_domino_heightfunc_to_tiling(hf, len(hf), d, len(d)))

    @staticmethod
    def TilingToHeightfunc(t, d):
        t, d = _i32(t), _i32(d)
        return _take_vec(lib().rt_domino_tiling_to_heightfunc(t, len(t), d, len(d)))

    @staticmethod
    def TilingToDomain(t):
        t = _i32(t)
        return _take_vec(lib().rt_domino_tiling_to_domain(t, len(t)))

    @staticmethod
    def DomainToVertices(d):
        d = _i32(d)
        return _take_vec(lib().rt_domino_domain_to_vertices(d, len(d)))

    @staticmethod
    def DomainToSVG(d, path):
        _svg("rt_domino_domain_svg", d, path)

    @staticmethod
    def TilingToSVG(t, path):
        _svg("rt_domino_tiling_svg", t, path)

    @staticmethod
    def MayaToSVG(t, path):
        _svg("rt_domino_maya_svg", t, path)

    @staticmethod
    def DimerToSVG(t, path, d=None):
        _dimer_svg("rt_domino_dimer_svg", t, d, path)


class LozengeTiler:
    """Static helpers of the reference's LozengeTiler (LozengeTiler.h:77-92)."""

    @staticmethod
    def AlmostHexDomain(n, m):
        return _take_vec(lib().rt_lozenge_almost_hex_domain(n, m))

    @staticmethod
    def MaxTiling(d):
        return _thurston("rt_lozenge_max_tiling", d)

    @staticmethod
    def MinTiling(d):
        return _thurston("rt_lozenge_min_tiling", d)

    @staticmethod
    def HeightfuncToTiling(hf, d):
        hf, d = _i32(hf), _i32(d)
        return _take_vec(lib().rt_lozenge_heightfunc_to_tiling(hf, len(hf), d, len(d)))

    @staticmethod
    def TilingToHeightfunc(t, d):
        t, d = _i32(t), _i32(d)
        return _take_vec(lib().rt_lozenge_tiling_to_heightfunc(t, len(t), d, len(d)))

    @staticmethod
    def TilingToDomain(t):
        t = _i32(t)
        return _take_vec(lib().rt_lozenge_tiling_to_domain(t, len(t)))

    @staticmethod
    def DomainToVertices(d):
        d = _i32(d)
        return _take_vec(lib().rt_lozenge_domain_to_vertices(d, len(d)))

    @staticmethod
    def DomainToSVG(d, path):
        _svg("rt_lozenge_domain_svg", d, path)

    @staticmethod
    def TilingToSVG(t, path):
        _svg("rt_lozenge_tiling_svg", t, path)

    @staticmethod
    def DimerToSVG(t, path, d=None):
        _dimer_svg("rt_lozenge_dimer_svg", t, d, path)


class RectTriangleTiler:
    """Static helpers of the reference's RectTriangleTiler (RectTriangleTiler.h:77-82)."""

    @staticmethod
    def maxHex(n):
        return _take_vec(lib().rt_recttri_max_hex(n))

    @staticmethod
    def minHex(n):
        return _take_vec(lib().rt_recttri_min_hex(n))

    @staticmethod
    def slopeHex(n):
        return _take_vec(lib().rt_recttri_slope_hex(n))

    @staticmethod
    def TilingToSVG(t, path):
        _svg("rt_recttri_tiling_svg", t, path)


class TriangleDimerTiler:
    """Static helpers of the reference's TriangleDimerTiler (bibone model,
    TriangleDimerTiler.h:194-209)."""

    @staticmethod
    def IceCreamCone(N, M):
        return _take_vec(lib().rt_bibone_ice_cream_cone(N, M))

    @staticmethod
    def TestButterfly(kind):
        return _take_vec(lib().rt_bibone_test_butterfly(kind))

    @staticmethod
    def Hexagon(a, b, c):
        """Hexagon with sides a, b, c, a, b, c and a dimer cover (no
        reference counterpart); RTError when its vertex count is odd."""
        return _take_vec(lib().rt_bibone_hexagon(a, b, c))

    @staticmethod
    def TilingToTriangleFlips(t):
        t = _i32(t)
        return _take_vec(lib().rt_bibone_triangle_flips(t, len(t)))

    @staticmethod
    def TilingToButterflyFlips(t):
        t = _i32(t)
        return _take_vec(lib().rt_bibone_butterfly_flips(t, len(t)))

    @staticmethod
    def TilingToVertices(t):
        t = _i32(t)
        return _take_vec(lib().rt_bibone_tiling_to_vertices(t, len(t)))

    @staticmethod
    def TilingToDomain(t):
        t = _i32(t)
        return _take_vec(lib().rt_bibone_tiling_to_domain(t, len(t)))

    @staticmethod
    def DimerIndicator(t, c):
        t = _i32(t)
        return _take_vec(lib().rt_bibone_dimer_indicator(t, len(t), c))

    @staticmethod
    def EdgeForm(t):
        """One byte per plane cell (bit 0 h, 1 l, 2 r dimer edges), as
        rt_bibone_walk runs the start on its edge form; None when the start
        runs on the eight planes instead (no reference counterpart)."""
        t = _i32(t)
        M = side(t) // 2
        out = np.zeros(M * M, dtype=np.uint8)
        rc = lib().rt_bibone_edge_form(t, len(t), out.ctypes.data_as(C.c_void_p))
        return out.reshape(M, M) if rc == 0 else None

    @staticmethod
    def DomainToSVG(d, path):
        _svg("rt_bibone_domain_svg", d, path)

    @staticmethod
    def TilingToSVG(t, path):
        _svg("rt_bibone_tiling_svg", t, path)

    @staticmethod
    def DimerToSVG(t, path, d=None):
        _dimer_svg("rt_bibone_dimer_svg", t, d, path)


def synthetic_params(count, seed=0x5eed):
    """Synthetic TinyMT parameter table (the reference's 2^20-line table is
    not in its checkout): per id a (mat1, mat2, tmat) triple from a
    splitmix64 stream of `seed`; seed=None gives the TINYMT32J constants for
    every id (tinymt32def.clh:48-50)."""
    if seed is None:
        p = np.empty((count, 3), dtype=np.uint32)
        p[:, 0], p[:, 1], p[:, 2] = 0x8F7011EE, 0xFC78FF1F, 0x3793FDFF
        return p
    idx = np.arange(3 * count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z & np.uint64(0xFFFFFFFF)).astype(np.uint32).reshape(count, 3)


def write_params_file(path, params3):
    """Write a table in the tinymt32dc text layout the reference parses."""
    with open(path, "w") as f:
        f.write("# synthetic TinyMT32 parameter table\n")
        for k, (m1, m2, tm) in enumerate(np.asarray(params3, dtype=np.uint32)):
            f.write(f"synthetic,127,{k},{int(m1):08x},{int(m2):08x},{int(tm):08x},0,0\n")

"""ctypes view of oracle/liboracle.so -- the CPU restatement used as the
checker.  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_L = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_LONGP = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class Tiny(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("s0", "s1", "s2", "s3", "mat1", "mat2", "tmat")]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


class Minstd(C.Structure):
    _fields_ = [("x", C.c_uint64)]


class CftpStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("total_steps", C.c_long), ("steps_run", C.c_long)]


_L.orc_tinymt_init.argtypes = [C.POINTER(Tiny), C.c_uint32]
_L.orc_tinymt_uint32.argtypes = [C.POINTER(Tiny)]
_L.orc_tinymt_uint32.restype = C.c_uint32
_L.orc_tinymt_single01.argtypes = [C.POINTER(Tiny)]
_L.orc_tinymt_single01.restype = C.c_float
_L.orc_tinymt_h_float01.argtypes = [C.POINTER(Tiny)]
_L.orc_tinymt_h_float01.restype = C.c_float
_L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
_L.orc_mt_next.argtypes = [C.POINTER(MT)]
_L.orc_mt_next.restype = C.c_uint32
_L.orc_minstd_seed.argtypes = [C.POINTER(Minstd), C.c_ulong]
_L.orc_minstd_next.argtypes = [C.POINTER(Minstd)]
_L.orc_minstd_next.restype = C.c_uint32
_L.orc_uniform_int.argtypes = [C.POINTER(Minstd), C.c_int, C.c_int]
_L.orc_parse_params.argtypes = [C.c_char_p, C.c_size_t, _U32P]
_L.orc_domino_walk.argtypes = [_I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_size_t, C.c_int,
                               C.c_double, C.c_double]
_L.orc_domino_walk_segments.argtypes = [_I32P, C.c_int, _LONGP, _LONGP, C.c_int, _U32P,
                                        C.c_size_t, C.c_int, C.c_double, C.c_double]
_L.orc_domino_cftp.argtypes = [_I32P, _I32P, C.c_int, C.c_uint32, C.c_int, C.c_long, C.c_int,
                               _U32P, C.c_size_t, C.c_int, C.c_double, C.c_double, _I32P,
                               C.POINTER(CftpStats)]
_L.orc_lozenge_walk.argtypes = [_I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_size_t]
_L.orc_lozenge_walk_segments.argtypes = [_I32P, C.c_int, _LONGP, _LONGP, C.c_int, _U32P,
                                         C.c_size_t]
_L.orc_lozenge_cftp.argtypes = [_I32P, _I32P, C.c_int, C.c_uint32, C.c_int, C.c_long, C.c_int,
                                _U32P, C.c_size_t, _I32P, C.POINTER(CftpStats)]
_L.orc_recttri_walk.argtypes = [_I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_size_t, _F32P]
_L.orc_bibone_walk.argtypes = [_I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_size_t,
                               C.c_void_p]
_L.orc_bibone_triangle_flips.argtypes = [_I32P, C.c_int, _I32P]
_L.orc_bibone_triangle_flips.restype = None
_L.orc_bibone_butterfly_flips.argtypes = [_I32P, C.c_int, _I32P]
_L.orc_bibone_butterfly_flips.restype = None
_L.orc_recttri_cftp.argtypes = [_I32P, _I32P, C.c_int, C.c_uint32, C.c_int, C.c_long, C.c_int,
                                _U32P, C.c_size_t, _F32P, _I32P, C.c_void_p]
_L.orc_recttri_walk_segments.argtypes = [_I32P, C.c_int, _LONGP, _LONGP, C.c_int, _U32P,
                                         C.c_size_t, _F32P]
_L.orc_set_threads.argtypes = [C.c_int]

set_threads = _L.orc_set_threads


def _p(params3):
    return np.ascontiguousarray(np.asarray(params3, dtype=np.uint32).ravel())


def _side(t):
    n = int(round(np.sqrt(len(t))))
    assert n * n == len(t)
    return n


def tiny(mat1, mat2, tmat, seed):
    t = Tiny(0, 0, 0, 0, mat1, mat2, tmat)
    _L.orc_tinymt_init(C.byref(t), seed)
    return t


def tiny_floats(t, n, h=False):
    f = _L.orc_tinymt_h_float01 if h else _L.orc_tinymt_single01
    return np.array([f(C.byref(t)) for _ in range(n)], dtype=np.float32)


def mt19937(seed, n):
    g = MT()
    _L.orc_mt_seed(C.byref(g), seed)
    return np.array([_L.orc_mt_next(C.byref(g)) for _ in range(n)], dtype=np.uint64)


def minstd(seed, n):
    g = Minstd()
    _L.orc_minstd_seed(C.byref(g), seed)
    return np.array([_L.orc_minstd_next(C.byref(g)) for _ in range(n)], dtype=np.uint64)


def uniform012(seed, n):
    g = Minstd()
    _L.orc_minstd_seed(C.byref(g), seed)
    return np.array([_L.orc_uniform_int(C.byref(g), 0, 2) for _ in range(n)], dtype=np.int32)


def parse_params(path, count):
    out = np.zeros(3 * count, dtype=np.uint32)
    rc = _L.orc_parse_params(str(path).encode(), count, out)
    return rc, out.reshape(count, 3)


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


def domino_walk(t, steps, seed, params3, rule=0, a=1.0, b=1.0):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    _check(_L.orc_domino_walk(t, _side(t), steps, seed, p, len(p) // 3, rule, a, b))
    return t


def domino_walk_segments(t, steps, seeds, params3, rule=0, a=1.0, b=1.0):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    st = np.ascontiguousarray(steps, dtype=np.int64)
    se = np.ascontiguousarray(seeds, dtype=np.int64)
    _check(_L.orc_domino_walk_segments(t, _side(t), st, se, len(st), p, len(p) // 3, rule, a, b))
    return t


def domino_cftp(tmax, tmin, master, params3, sched=0, t0=128, max_rounds=40, rule=1, a=1.0,
                b=1.0):
    tmax = np.ascontiguousarray(tmax, dtype=np.int32)
    tmin = np.ascontiguousarray(tmin, dtype=np.int32)
    out = np.zeros_like(tmax)
    st = CftpStats()
    p = _p(params3)
    rc = _L.orc_domino_cftp(tmax, tmin, _side(tmax), master, sched, t0, max_rounds, p,
                            len(p) // 3, rule, a, b, out, C.byref(st))
    _check(rc)
    return out, {"rounds": st.rounds, "total_steps": st.total_steps, "steps_run": st.steps_run}


def lozenge_walk(t, steps, seed, params3):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    _check(_L.orc_lozenge_walk(t, _side(t), steps, seed, p, len(p) // 3))
    return t


def lozenge_walk_segments(t, steps, seeds, params3):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    st = np.ascontiguousarray(steps, dtype=np.int64)
    se = np.ascontiguousarray(seeds, dtype=np.int64)
    _check(_L.orc_lozenge_walk_segments(t, _side(t), st, se, len(st), p, len(p) // 3))
    return t


def lozenge_cftp(tmax, tmin, master, params3, sched=0, t0=128, max_rounds=40):
    tmax = np.ascontiguousarray(tmax, dtype=np.int32)
    tmin = np.ascontiguousarray(tmin, dtype=np.int32)
    out = np.zeros_like(tmax)
    st = CftpStats()
    p = _p(params3)
    rc = _L.orc_lozenge_cftp(tmax, tmin, _side(tmax), master, sched, t0, max_rounds, p,
                             len(p) // 3, out, C.byref(st))
    _check(rc)
    return out, {"rounds": st.rounds, "total_steps": st.total_steps, "steps_run": st.steps_run}


def recttri_walk(t, steps, seed, params3, weights16=None):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    w = np.ones(16, dtype=np.float32) if weights16 is None else np.asarray(weights16, np.float32)
    if weights16 is None:
        w[0] = 0.0
    _check(_L.orc_recttri_walk(t, _side(t), steps, seed, p, len(p) // 3,
                               np.ascontiguousarray(w)))
    return t


def bibone_triangle_flips(t):
    t = np.ascontiguousarray(t, dtype=np.int32)
    N = _side(t)
    out = np.zeros((N // 2) * N, np.int32)
    _L.orc_bibone_triangle_flips(t, N, out)
    return out


def bibone_butterfly_flips(t):
    t = np.ascontiguousarray(t, dtype=np.int32)
    N = _side(t)
    out = np.zeros(N * N, np.int32)
    _L.orc_bibone_butterfly_flips(t, N, out)
    return out


def bibone_walk(t, steps, seed, params3, planes=False):
    """Returns the tiling, or (tiling, 8 x M x M planes) with planes=True."""
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = _side(t)
    M = N // 2
    p = _p(params3)
    pl = np.zeros(8 * M * M, np.int32) if planes else None
    _check(_L.orc_bibone_walk(t, N, steps, seed, p, len(p) // 3,
                              pl.ctypes.data_as(C.c_void_p) if planes else None))
    return (t, pl.reshape(8, M, M)) if planes else t


def recttri_cftp(tmax, tmin, master, params3, sched=0, t0=128, max_rounds=40, weights16=None):
    tmax = np.ascontiguousarray(tmax, dtype=np.int32)
    tmin = np.ascontiguousarray(tmin, dtype=np.int32)
    out = np.zeros_like(tmax)
    st = CftpStats()
    p = _p(params3)
    w = np.ones(16, dtype=np.float32) if weights16 is None else np.asarray(weights16, np.float32)
    if weights16 is None:
        w[0] = 0.0
    rc = _L.orc_recttri_cftp(tmax, tmin, _side(tmax), master, sched, t0, max_rounds, p,
                             len(p) // 3, np.ascontiguousarray(w), out, C.addressof(st))
    _check(rc)
    return out, {"rounds": st.rounds, "total_steps": st.total_steps, "steps_run": st.steps_run}


def recttri_walk_segments(t, steps, seeds, params3, weights16=None):
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    p = _p(params3)
    w = np.ones(16, dtype=np.float32) if weights16 is None else np.asarray(weights16, np.float32)
    if weights16 is None:
        w[0] = 0.0
    st = np.ascontiguousarray(steps, dtype=np.int64)
    se = np.ascontiguousarray(seeds, dtype=np.int64)
    _check(_L.orc_recttri_walk_segments(t, _side(t), st, se, len(st), p, len(p) // 3,
                                        np.ascontiguousarray(w)))
    return t

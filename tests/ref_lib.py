"""ctypes view of oracle/_ref/libref_host.so: the reference's own TinyMT code
(tinymt32.clh through clang's OpenCL C front end, tinymt32.h) and its
parameter-file reader (file_reader.cpp), built by `make -C oracle ref` from
the reference checkout.  Test infrastructure only; `available()` is False
where the reference checkout was not there to build it."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")

_U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_R = None


def available():
    return os.path.exists(PATH)


def _lib():
    global _R
    if _R is None:
        R = C.CDLL(PATH)
        R.ref_cl_tinymt_init.argtypes = [_U32P, C.c_uint32]
        R.ref_cl_tinymt_init.restype = None
        R.ref_cl_tinymt_single01.argtypes = [_U32P, _F32P, C.c_int]
        R.ref_cl_tinymt_single01.restype = None
        R.ref_cl_tinymt_uint32.argtypes = [_U32P, _U32P, C.c_int]
        R.ref_cl_tinymt_uint32.restype = None
        R.ref_h_tinymt_float01.argtypes = [_U32P, _F32P, C.c_int]
        R.ref_h_tinymt_float01.restype = None
        R.ref_h_tinymt_uint32.argtypes = [_U32P, _U32P, C.c_int]
        R.ref_h_tinymt_uint32.restype = None
        R.ref_parse_params.argtypes = [C.c_char_p, C.c_long, _U32P]
        R.ref_parse_params.restype = C.c_int
        _R = R
    return _R


def cl_init(mat1, mat2, tmat, seed):
    """tinymt32_init of tinymt32.clh: the state (s0..s3) for the params and seed."""
    s7 = np.array([0, 0, 0, 0, mat1, mat2, tmat], np.uint32)
    _lib().ref_cl_tinymt_init(s7, seed)
    return s7


def cl_single01(s7, n):
    """n tinymt32_single01 draws of tinymt32.clh from state s7 (advanced in place)."""
    out = np.zeros(n, np.float32)
    _lib().ref_cl_tinymt_single01(s7, out, n)
    return out


def cl_uint32(s7, n):
    out = np.zeros(n, np.uint32)
    _lib().ref_cl_tinymt_uint32(s7, out, n)
    return out


def h_float01(s7, n):
    """n tinymt32_generate_float01 draws of tinymt32.h from state s7."""
    out = np.zeros(n, np.float32)
    _lib().ref_h_tinymt_float01(s7, out, n)
    return out


def h_uint32(s7, n):
    out = np.zeros(n, np.uint32)
    _lib().ref_h_tinymt_uint32(s7, out, n)
    return out


def parse_params(path, count):
    """count file_reader::get calls: (rc, params[count, 3]); rc -2 where the
    reader throws, -3 where the file does not open."""
    out = np.zeros(3 * max(count, 1), np.uint32)
    rc = _lib().ref_parse_params(str(path).encode(), count, out)
    return rc, out[:3 * count].reshape(count, 3)

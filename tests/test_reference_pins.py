"""Pins the oracle's TinyMT generator, its InitTinyMT and its parameter-file
parser against the reference's own code, built from the reference sources
where they lie (oracle/Makefile `ref`):

* tinymt32.clh (the generator and tinymt32_init the kernels use,
  tinymt32.clh:39-55, 62-129, 163-203), compiled for the host by clang's
  OpenCL C front end (oracle/ref_tinymt_probe.cl);
* tinymt32.h (the host-side formulation, tinymt32.h:72-136, 205-208);
* file_reader::get (file_reader.cpp:51-70, 103-166), as get_params_buffer
  calls it (22-38).

Where oracle/_ref is not built (no reference checkout), the same checks run
against fixtures that tests/golden/make_golden.py wrote from that library
(tests/golden/reference_tinymt.npz, tests/golden/reference_parser.json).
"""
import json
import os

import numpy as np
import pytest

import randomtilingspublic_amd as rt
from tests import oracle_lib as o
from tests import ref_lib as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")
M32 = 0xFFFFFFFF


def _orc_init(mat1, mat2, tmat, seed):
    t = o.tiny(int(mat1), int(mat2), int(tmat), int(seed))
    return np.array([t.s0, t.s1, t.s2, t.s3], np.uint32)


def _orc_draws(s7, n):
    t = o.Tiny(*[int(x) for x in s7])
    f = np.array([o._L.orc_tinymt_single01(o.C.byref(t)) for _ in range(n)], np.float32)
    return f


def _orc_uint32(s7, n):
    t = o.Tiny(*[int(x) for x in s7])
    return np.array([o._L.orc_tinymt_uint32(o.C.byref(t)) for _ in range(n)], np.uint32)


def certification_params(final):
    """(seed, mat1, mat2, tmat) whose tinymt32_init mixing rounds
    (tinymt32.clh:188-195) end in the state `final`.  Round i xors
    st[i & 3] with a function of st[(i-1) & 3], which no later round has
    changed yet, so running the rounds backwards inverts them."""
    st = [int(v) & M32 for v in final]
    for i in range(7, 0, -1):
        p = st[(i - 1) & 3]
        st[i & 3] ^= (i + 1812433253 * (p ^ (p >> 30))) & M32
    return st  # [seed, mat1, mat2, tmat]


def _mix(seed, mat1, mat2, tmat):
    st = [seed, mat1, mat2, tmat]
    for i in range(1, 8):
        p = st[(i - 1) & 3]
        st[i & 3] ^= (i + 1812433253 * (p ^ (p >> 30))) & M32
    return st


def test_certification_inverse_is_exact():
    for final in ([0, 0, 0, 0], [0x80000000, 0, 0, 0], [5, 6, 7, 8]):
        assert _mix(*certification_params(final)) == final


def test_period_certification_branch():
    """A state whose mixed words are zero (bit 31 of s0 aside) is replaced by
    'T','I','N','Y' before the 8 warm-up steps (tinymt32.clh:163-174,
    197-202): the oracle's init equals 8 steps from that state."""
    for final in ([0, 0, 0, 0], [0x80000000, 0, 0, 0]):
        seed, m1, m2, tm = certification_params(final)
        got = _orc_init(m1, m2, tm, seed)
        t = o.Tiny(ord("T"), ord("I"), ord("N"), ord("Y"), m1, m2, tm)
        for _ in range(8):
            o._L.orc_tinymt_uint32(o.C.byref(t))
        assert got.tolist() == [t.s0, t.s1, t.s2, t.s3]
    # one bit off the zero state: no certification
    seed, m1, m2, tm = certification_params([1, 0, 0, 0])
    t = o.Tiny(1, 0, 0, 0, m1, m2, tm)
    for _ in range(8):
        o._L.orc_tinymt_uint32(o.C.byref(t))
    assert _orc_init(m1, m2, tm, seed).tolist() == [t.s0, t.s1, t.s2, t.s3]


def _random_cases(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint64).astype(np.uint32)


@needs_ref
def test_tinymt_init_matches_reference_opencl_code():
    cases = _random_cases(10000, 1)
    extra = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [1, 0x8F7011EE, 0xFC78FF1F, 0x3793FDFF]]
    for final in ([0, 0, 0, 0], [0x80000000, 0, 0, 0]):
        extra.append(certification_params(final))
    for seed, m1, m2, tm in [c[:4] for c in cases] + [np.array(e, np.uint64) for e in extra]:
        want = R.cl_init(m1, m2, tm, seed)
        assert np.array_equal(_orc_init(m1, m2, tm, seed), want[:4]), (seed, m1, m2, tm)


@needs_ref
def test_tinymt_draws_match_reference_opencl_code():
    for c in _random_cases(10000, 2):
        s7 = np.ascontiguousarray(c[:7])
        a, b = s7.copy(), s7.copy()
        want = R.cl_single01(a, 8)
        assert np.array_equal(_orc_draws(s7, 8).view(np.uint32), want.view(np.uint32))
        assert np.array_equal(_orc_uint32(s7, 8), R.cl_uint32(b, 8))


@needs_ref
def test_tinymt_draws_match_reference_host_header():
    for c in _random_cases(10000, 3):
        s7 = np.ascontiguousarray(c[:7])
        a, b = s7.copy(), s7.copy()
        want = R.h_float01(a, 8)
        assert np.array_equal(_orc_draws(s7, 8).view(np.uint32), want.view(np.uint32))
        assert np.array_equal(_orc_uint32(s7, 8), R.h_uint32(b, 8))


def test_tinymt_golden_vectors():
    """Fixture written by the reference's tinymt32.clh (make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "reference_tinymt.npz"))
    for (seed, m1, m2, tm), st in zip(g["init_in"], g["init_out"]):
        assert np.array_equal(_orc_init(m1, m2, tm, seed), st)
    for s7, f, u in zip(g["draw_in"], g["draw_f32"], g["draw_u32"]):
        assert np.array_equal(_orc_draws(s7, f.size).view(np.uint32), f.view(np.uint32))
        assert np.array_equal(_orc_uint32(s7, u.size), u)


def parser_cases():
    """(name, file bytes, count): the layouts the tinymt32dc reader meets."""
    p = rt.synthetic_params(6, 1234)
    lines = [f"tinymt32dc,127,{k},{int(a):08x},{int(b):08x},{int(c):08x},63,0"
             for k, (a, b, c) in enumerate(p)]
    nl = "\n".join(lines) + "\n"
    return [
        ("plain", nl, 6),
        ("prefix", nl, 3),
        ("header_and_comments", "# header\n#\n" + "\n# mid\n".join(lines) + "\n", 6),
        ("crlf", "\r\n".join(lines) + "\r\n", 6),
        ("no_final_newline", "\n".join(lines), 6),
        ("short_file", nl, 7),
        ("extra_fields", "\n".join(l + ",9,9,9" for l in lines) + "\n", 6),
        ("only_five_commas", "a,127,0,1f,2E,3d\nb,127,1,0x10,0X20,30\n", 2),
        ("upper_hex_and_spaces", "n,1,2, 8F7011EE , FC78FF1F,3793FDFF\n", 1),
        ("wider_than_32_bits", "n,1,2,1ffffffff,2,3\n", 1),
        ("past_unsigned_long", "n,1,2,1ffffffffffffffff,2,3\n", 1),
        ("four_commas", "n,127,0,8f7011ee,fc78ff1f\n", 1),
        ("blank_line", lines[0] + "\n\n" + lines[1] + "\n", 2),
        ("cr_blank_line", lines[0] + "\r\n\r\n" + lines[1] + "\r\n", 2),
        ("empty_file", "", 1),
        ("zero_count", nl, 0),
    ]


def _write(tmp_path, name, text):
    path = tmp_path / f"{name}.txt"
    path.write_bytes(text.encode())
    return path


def _engine_parse(path, count):
    try:
        return 0, rt.parse_tinymt_file(path, count).reshape(count, 3)
    except rt.RTError as e:
        return e.code, None


def _agree(tmp_path, name, text, count, want_rc, want):
    path = _write(tmp_path, name, text)
    rc, got = o.parse_params(path, count)
    assert rc == want_rc, name
    erc, eng = _engine_parse(path, count)
    assert erc == want_rc, name
    if want_rc == 0:
        assert np.array_equal(got, want), name
        assert np.array_equal(eng, want), name


@needs_ref
def test_param_parser_matches_reference_reader(tmp_path):
    for name, text, count in parser_cases():
        rc, want = R.parse_params(_write(tmp_path, "ref_" + name, text), count)
        _agree(tmp_path, name, text, count, rc, want)
    assert R.parse_params(tmp_path / "missing.txt", 1)[0] == -3


def test_param_parser_golden(tmp_path):
    """Fixture written by the reference's file_reader (make_golden.py)."""
    with open(os.path.join(GOLDEN, "reference_parser.json")) as f:
        g = json.load(f)
    assert {c["name"] for c in g} == {c[0] for c in parser_cases()}
    for c in g:
        want = np.array(c["params"], np.uint32).reshape(c["count"], 3)
        _agree(tmp_path, c["name"], c["text"], c["count"], c["rc"], want)

"""Regenerates the golden fixtures under tests/golden/.

* schedules.json: colour schedules from libstdc++'s own mt19937 and
  default_random_engine + uniform_int_distribution (schedule_golden.cpp,
  compiled with the image's g++).
* reference_host.npz: outputs of the reference's own host statics
  (MinTiling / MaxTiling / AlmostHexDomain / maxHex / minHex ...), taken
  from oracle/_ref/libref_host.so, which oracle/Makefile builds from the
  reference sources.  Skipped when that library is not built.
* reference_outputs/*.txt: SaveMatrix files the reference checkout ships
  in Examples/ExampleOuts (copied data, not code).
* reference_tinymt.npz: tinymt32_init states and tinymt32_single01 /
  tinymt32_uint32 draws of the reference's own tinymt32.clh (compiled for
  the host into oracle/_ref/libref_host.so) on random parameters, seeds and
  states.
* reference_parser.json: what the reference's file_reader returns (or
  that it throws) on the parameter-file layouts of
  tests/test_reference_pins.py.
"""
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def schedules():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "sched")
        subprocess.run(["g++", "-O2", "-std=c++11", "-o", exe,
                        os.path.join(HERE, "schedule_golden.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    with open(os.path.join(HERE, "schedules.json"), "w") as f:
        json.dump(data, f)


def ref_host():
    lib = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")
    if not os.path.exists(lib):
        print("oracle/_ref/libref_host.so not built; skipping reference_host.npz")
        return
    R = C.CDLL(lib)

    def call(fn, *args):
        cap = 1 << 22
        out = np.zeros(cap, np.int32)
        f = getattr(R, fn)
        f.restype = C.c_long
        conv = []
        for a in args:
            if isinstance(a, np.ndarray):
                conv += [a.ctypes.data_as(C.c_void_p), C.c_long(len(a))]
            else:
                conv.append(C.c_int(a))
        n = f(*conv, out.ctypes.data_as(C.c_void_p), C.c_long(cap))
        assert n >= 0, fn
        return out[:n].copy()

    arrs = {}
    for n in (1, 2, 3, 7, 16, 33):
        d = call("ref_domino_aztec_diamond", n)
        arrs[f"domino_aztec{n}_domain"] = d
        arrs[f"domino_aztec{n}_min"] = call("ref_domino_min_tiling", d)
        arrs[f"domino_aztec{n}_max"] = call("ref_domino_max_tiling", d)
        arrs[f"domino_aztec{n}_min_hf"] = call("ref_domino_tiling_to_heightfunc",
                                               arrs[f"domino_aztec{n}_min"], d)
    rem = np.array([0, 2], np.int32)
    d = call("ref_domino_aztec_rectangle", 6, rem)  # (ptr, k) from the array
    arrs["domino_aztecrect6_02_domain"] = d
    arrs["domino_aztecrect6_02_min"] = call("ref_domino_min_tiling", d)
    for m, n in ((4, 6), (7, 10)):
        d = call("ref_domino_rectangle", m, n)
        arrs[f"domino_rect{m}x{n}_domain"] = d
        arrs[f"domino_rect{m}x{n}_max"] = call("ref_domino_max_tiling", d)
    for n, m in ((2, 0), (5, 1), (10, 2), (13, 0)):
        d = call("ref_lozenge_almost_hex_domain", n, m)
        arrs[f"lozenge_hex{n}_{m}_domain"] = d
        arrs[f"lozenge_hex{n}_{m}_min"] = call("ref_lozenge_min_tiling", d)
        arrs[f"lozenge_hex{n}_{m}_max"] = call("ref_lozenge_max_tiling", d)
        arrs[f"lozenge_hex{n}_{m}_max_hf"] = call("ref_lozenge_tiling_to_heightfunc",
                                                  arrs[f"lozenge_hex{n}_{m}_max"], d)
    for n in (1, 2, 5, 8):
        arrs[f"recttri_max{n}"] = call("ref_recttri_max_hex", n)
        arrs[f"recttri_min{n}"] = call("ref_recttri_min_hex", n)
        arrs[f"recttri_slope{n}"] = call("ref_recttri_slope_hex", n)
    for n, m in ((3, 0), (6, 0), (7, 2), (10, 3), (13, 5)):
        t = call("ref_bibone_ice_cream_cone", n, m)
        arrs[f"bibone_cone{n}_{m}"] = t
        arrs[f"bibone_cone{n}_{m}_tri"] = call("ref_bibone_triangle_flips", t)
        arrs[f"bibone_cone{n}_{m}_bf"] = call("ref_bibone_butterfly_flips", t)
        arrs[f"bibone_cone{n}_{m}_vertices"] = call("ref_bibone_tiling_to_vertices", t)
        arrs[f"bibone_cone{n}_{m}_domain"] = call("ref_bibone_tiling_to_domain", t)
        for c in range(3):
            arrs[f"bibone_cone{n}_{m}_dimers{c}"] = call("ref_bibone_dimer_indicator", t, c)
    for k in range(3):
        arrs[f"bibone_butterfly{k}"] = call("ref_bibone_test_butterfly", k)
    np.savez_compressed(os.path.join(HERE, "reference_host.npz"), **arrs)


def ref_tinymt():
    sys.path.insert(0, ROOT)
    from tests import ref_lib as R
    if not R.available():
        print("oracle/_ref/libref_host.so not built; skipping reference_tinymt.npz")
        return
    rng = np.random.default_rng(20240601)
    init_in = rng.integers(0, 1 << 32, size=(400, 4), dtype=np.uint64).astype(np.uint32)
    init_out = np.stack([R.cl_init(m1, m2, tm, sd)[:4] for sd, m1, m2, tm in init_in])
    draw_in = rng.integers(0, 1 << 32, size=(400, 7), dtype=np.uint64).astype(np.uint32)
    f, u = [], []
    for s7 in draw_in:
        f.append(R.cl_single01(s7.copy(), 8))
        u.append(R.cl_uint32(s7.copy(), 8))
    np.savez_compressed(os.path.join(HERE, "reference_tinymt.npz"), init_in=init_in,
                        init_out=init_out, draw_in=draw_in, draw_f32=np.stack(f),
                        draw_u32=np.stack(u))


def ref_parser():
    sys.path.insert(0, ROOT)
    from tests import ref_lib as R
    from tests.test_reference_pins import parser_cases
    if not R.available():
        print("oracle/_ref/libref_host.so not built; skipping reference_parser.json")
        return
    out = []
    with tempfile.TemporaryDirectory() as d:
        for name, text, count in parser_cases():
            path = os.path.join(d, name + ".txt")
            with open(path, "wb") as f:
                f.write(text.encode())
            rc, p = R.parse_params(path, count)
            out.append({"name": name, "text": text, "count": count, "rc": rc,
                        "params": p.ravel().tolist() if rc == 0 else [0] * (3 * count)})
    with open(os.path.join(HERE, "reference_parser.json"), "w") as f:
        json.dump(out, f, indent=1)


def svg_digests():
    """SHA-256 of the reference's example SVGs whose inputs are deterministic
    (Examples/ExampleOuts), keyed by the call that made them.  One line of
    TriangleDimer/DimerStart.svg comes from an out-of-bounds read in the
    reference (TriangleDimerTiler.cpp:1080-1082 reads d[(i-1)*2*M+j] at
    i = 0, a negative index, which happened to hold 1); that line is dropped
    before hashing, and the engine's writer reads outside the domain as 0."""
    import hashlib
    out_dir = "/root/reference/Examples/ExampleOuts"
    if not os.path.isdir(out_dir):
        print("reference checkout absent; skipping svg_digests.json")
        return
    files = {
        "domino_tiling_MinimalDomino_StartTiling": "MinimalDomino/StartTiling.svg",
        "domino_tiling_aztec16_min": "WeightedAztecDiamond/TilingStart.svg",
        "domino_tiling_aztec16_random_end": "WeightedAztecDiamond/TilingEnd.svg",
        "lozenge_domain_hex10_2": "MinimalLozenge/domainAlmostHex.svg",
        "lozenge_tiling_hex10_2_max": "MinimalLozenge/tilingStartAlmostHex.svg",
        "recttri_tiling_max5": "RectTriangle/MaxTiling.svg",
        "recttri_tiling_min5": "RectTriangle/MinTiling.svg",
        "recttri_tiling_slope5": "RectTriangle/SlopeTiling.svg",
        "recttri_tiling_random_end5": "RectTriangle/EndTiling.svg",
        "bibone_domain_cone6": "TriangleDimer/Domain.svg",
        "bibone_tiling_cone6": "TriangleDimer/TilingStart.svg",
        "bibone_dimer_cone6": "TriangleDimer/DimerStart.svg",
    }
    dig = {}
    for key, rel in files.items():
        data = open(os.path.join(out_dir, rel), "rb").read()
        if key == "bibone_dimer_cone6":
            data = data.replace(b'<use xlink:href = "#eH" x = "3.5" y = "0"/>\n', b"", 1)
        dig[key] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data), "file": rel}
    with open(os.path.join(HERE, "svg_digests.json"), "w") as f:
        json.dump(dig, f, indent=1)


if __name__ == "__main__":
    schedules()
    ref_host()
    ref_tinymt()
    ref_parser()
    svg_digests()
    print("golden fixtures written", file=sys.stderr)

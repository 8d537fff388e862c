"""The C-ABI library loads without a GPU and exports every entry point that
include/rtiling.h declares (names only: no compute call is made here)."""
import ctypes as C
import os
import re

import randomtilingspublic_amd as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "rtiling.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    assert len(names) > 60
    for must in ("rt_create", "rt_domino_walk", "rt_lozenge_cftp", "rt_bibone_walk",
                 "rt_domino_walk_batch_observe", "rt_server_submit", "rt_set_autotune"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(rt.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing

"""The C-ABI library loads and exports every entry point include/cpsfdtd.h
declares (no device calls)."""
import ctypes
import os
import re

import cpsfdtd_amd as eng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "cpsfdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 40
    assert set(names) == set(eng.SIGNATURES), set(names) ^ set(eng.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(eng.LIB_PATH)
    for n in declared():
        assert hasattr(lib, n), n
    eng.library()  # binds every signature


def test_library_is_gfx950_hip():
    data = open(eng.LIB_PATH, "rb").read()
    assert b"gfx950" in data

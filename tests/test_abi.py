"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/ves3d_hip.h declares; the Python binding covers exactly that set."""
import ctypes
import os
import re

import ves3d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\bint\s+(ves3d_\w+)\s*\(", txt))
    return names


def test_library_exports_declared_symbols():
    decl = _declared()
    assert len(decl) > 20
    lib = ctypes.CDLL(ves3d_amd.LIB_PATH)
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert set(ves3d_amd.exported_symbols()) == _declared()

"""The C-ABI library builds for gfx950 and exports every entry point that
include/tpls_core.h and include/tpls_io.h declare (no GPU needed: nothing
is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("tpls_core.h", "tpls_io.h")]
LIB = os.path.join(ROOT, "tpls_amd", "libtpls_hip.so")


def declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(tpls_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "tpls_step" in names and "tpls_pressure_solve" in names
    assert "tpls_load_data_into_array" in names and "tpls_load_options" in names
    assert len(names) >= 55


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_has_gfx950_code():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    blob = open(LIB, "rb").read()
    assert b"gfx950" in blob

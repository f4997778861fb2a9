"""The C-ABI library loads and exports every entry point include/quack_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "quack_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qk_\w+)\s*\(", text)))


def test_header_parses():
    names = declared()
    assert "qk_rg0w0" in names and "qk_evrgw" in names and len(names) >= 15


def test_library_exports_all_declared_symbols():
    from quack_amd import gw
    if not os.path.exists(gw.library_path()):
        pytest.fail("libquack_amd.so missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(gw.library_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gw.EXPORTS) == declared()


def test_no_cpu_fallback_in_host_mirror():
    """The host mirror never imports the oracle and fails loudly without the library."""
    src = open(os.path.join(ROOT, "quack_amd", "gw.py")).read()
    assert "oracle" not in src.split('"""', 2)[2]
    from quack_amd import gw
    saved = gw.LIB_PATH, gw._lib
    try:
        gw.LIB_PATH, gw._lib = "/nonexistent/libquack_amd.so", None
        with pytest.raises(RuntimeError):
            gw.load()
    finally:
        gw.LIB_PATH, gw._lib = saved

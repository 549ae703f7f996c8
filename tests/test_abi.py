"""The C-ABI library loads and exports every symbol include/wb97mv_amd.h
declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, 'include', 'wb97mv_amd.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|float|const char\*)\s+(wb_\w+)\s*\(', text, re.M)))


def test_header_and_binding_agree():
    from pyscf_wb97mv_fast_amd import lib
    assert sorted(lib.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from pyscf_wb97mv_fast_amd import lib
    so = ctypes.CDLL(lib.LIB_PATH)
    for name in _declared():
        assert hasattr(so, name), name
    lib.load()
    assert lib.load().wb_version() == 1

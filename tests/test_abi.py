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


def test_library_exports_nothing_else():
    """Internal functions stay hidden: the exported code symbols are the
    declared entry points and no library-internal (wb namespace) function.
    (The host-side handles of the device kernels are data objects that the
    HIP toolchain always exports; they are not functions.)"""
    import shutil
    import subprocess
    import pytest
    from pyscf_wb97mv_fast_amd import lib
    nm = shutil.which('nm')
    if nm is None:
        pytest.skip('nm not available')
    out = subprocess.run([nm, '-D', '--defined-only', lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines()
            if len(line.split()) == 3 and line.split()[1] in ('T', 'W', 'i')}
    ours = {s for s in syms if s.startswith('wb_') or s.startswith('_ZN2wb')}
    assert ours == set(_declared()), sorted(ours ^ set(_declared()))

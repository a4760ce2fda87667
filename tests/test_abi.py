"""The C-ABI library loads on the CPU host and exports every symbol include/mvd_rhs.h declares."""
import ctypes
import os
import subprocess

import pytest

from conftest import PKG


def test_library_exports_header_symbols(mvd):
    lib_path = mvd.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.fail("libmvd_rhs.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(lib_path)
    declared = mvd.exported_symbols_declared()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"
    nm = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True)
    exported = {line.split()[-1] for line in nm.stdout.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_create_rejects_bad_params_without_gpu(mvd):
    """Argument checks run before any device call."""
    import ctypes as C
    p = mvd.make_params(2, 8, 8)          # nx < 3 (SetupDecomp requires >= 3)
    h = C.c_void_p()
    assert mvd.lib().mvd_create(C.byref(p), None, C.byref(h)) == mvd.MVD_BAD_ARGUMENT
    p = mvd.make_params(8, 8, 8, bc=(0, 1, 0, 0, 0, 0))   # half-periodic x
    assert mvd.lib().mvd_create(C.byref(p), None, C.byref(h)) == mvd.MVD_BAD_ARGUMENT
    p = mvd.make_params(8, 8, 8, nchem=3)
    assert mvd.lib().mvd_create(C.byref(p), None, C.byref(h)) == mvd.MVD_UNSUPPORTED

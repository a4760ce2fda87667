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


def test_arkode_library_exports_header_symbols(mvd):
    """libmvd_arkode.so (the ARKODE callbacks) loads once SUNDIALS' symbols are resolvable
    (here the tests' stand-in) and exports every function include/mvd_arkode.h declares."""
    import arkode_harness
    arkode_harness.stub()
    L = mvd.arkode_lib()
    declared = mvd.exported_symbols_declared(mvd.ARKODE_HEADER_PATH)
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"
    nm = subprocess.run(["nm", "-D", "--defined-only", mvd.ARKODE_LIB_PATH], capture_output=True,
                        text=True)
    exported = {line.split()[-1] for line in nm.stdout.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_arkode_data_layout_matches_header(mvd):
    """The Python mirror of mvd_arkode_data has the header's field order and size: compile a
    probe against include/mvd_arkode.h and compare offsets."""
    import ctypes as C
    import tempfile
    fields = [f for f, _ in mvd.ArkodeData._fields_]
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"mvd_arkode.h\"\nint main(){"
    src += 'printf("%zu", sizeof(mvd_arkode_data));'
    for f in fields:
        src += f'printf(" %zu", offsetof(mvd_arkode_data, {f}));'
    src += "return 0;}\n"
    inc = os.path.join(os.path.dirname(PKG), "include")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "probe")
        subprocess.run(["gcc", "-I", inc, c, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    assert out[0] == C.sizeof(mvd.ArkodeData)
    assert out[1:] == [getattr(mvd.ArkodeData, f).offset for f in fields]

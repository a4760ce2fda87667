"""Test harness for the ARKODE callbacks (libmvd_arkode.so): loads the test-only SUNDIALS
stand-in (tests/sundials_stub) with RTLD_GLOBAL so that the adapters' SUNDIALS symbols
resolve, and builds stand-in ManyVectors and block matrices over torch device tensors."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
STUB_PATH = os.path.join(HERE, "sundials_stub", "libsundials_stub.so")
_stub = None


def stub():
    global _stub
    if _stub is not None:
        return _stub
    if not os.path.exists(STUB_PATH):
        raise RuntimeError("SUNDIALS stand-in not built: make -C tests/sundials_stub")
    L = C.CDLL(STUB_PATH, mode=C.RTLD_GLOBAL)
    vp, d = C.c_void_p, C.c_double
    sig = {
        "stub_vector": ([vp, C.c_long], vp),
        "stub_manyvector": ([C.c_int, C.POINTER(vp)], vp),
        "stub_vector_free": ([vp], None),
        "stub_matrix": ([vp, C.c_long], vp),
        "stub_matrix_free": ([vp], None),
        "stub_context": ([], vp),
        "stub_context_free": ([vp], None),
        "stub_set_forcing": ([vp, d, d, C.c_int, C.POINTER(vp)], None),
        "stub_stepper_evolve": ([vp, d, d, vp], C.c_int),
        "stub_stepper_fullrhs": ([vp, d, vp, vp, C.c_int], C.c_int),
        "stub_stepper_reset": ([vp, d, vp], C.c_int),
        "stub_ls_gettype": ([vp], C.c_int),
        "stub_ls_initialize": ([vp], C.c_int),
        "stub_ls_setup": ([vp, vp], C.c_int),
        "stub_ls_solve": ([vp, vp, vp, vp, d], C.c_int),
        "stub_ls_lastflag": ([vp], C.c_long),
        "stub_ls_free": ([vp], C.c_int),
        "MRIStepInnerStepper_Free": ([C.POINTER(vp)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _stub = L
    return L


class ManyVector:
    """A stand-in ManyVector over torch device tensors (kept alive with it)."""

    def __init__(self, tensors):
        L = stub()
        self.tensors = list(tensors)
        self.leaves = [L.stub_vector(C.c_void_p(t.data_ptr()), t.numel()) for t in self.tensors]
        arr = (C.c_void_p * len(self.leaves))(*self.leaves)
        self.h = L.stub_manyvector(len(self.leaves), arr)

    def free(self):
        L = stub()
        L.stub_vector_free(self.h)
        for v in self.leaves:
            L.stub_vector_free(v)


class Leaf:
    """A stand-in device vector (one array)."""

    def __init__(self, t):
        self.t = t
        self.h = stub().stub_vector(C.c_void_p(t.data_ptr()), t.numel())

    def free(self):
        stub().stub_vector_free(self.h)


class BlockMatrix:
    """A stand-in SUNMatrix_MagmaDenseBlock(nblocks, 10, 10) over a device tensor."""

    def __init__(self, t, nblocks):
        self.t = t
        self.h = stub().stub_matrix(C.c_void_p(t.data_ptr()), nblocks)

    def free(self):
        stub().stub_matrix_free(self.h)

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sundials-manyvector-demo_amd")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmvd_rhs.so)")


def load_pyapi():
    spec = importlib.util.spec_from_file_location("mvd_pyapi", os.path.join(PKG, "pyapi.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules.setdefault("mvd_pyapi", mod)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mvd():
    return load_pyapi()


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def normwise_err(a, b):
    """max |a-b| / max |b| (field-norm-relative error; fe and rates are ~0 in smooth regions)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)

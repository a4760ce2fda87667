"""fe() on the GPU through the C ABI, against the CPU oracle.

Tolerances:
  plain kernels (reference operation order, no FMA)  -> bitwise equal to the oracle
  fast kernels                                        -> field-norm-relative error <= 1e-12
"""
import importlib.util
import os
import sys

import numpy as np
import pytest

import orc
from conftest import PKG, normwise_err

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def to_dev(torch, arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in arrs]


def zeros_dev(torch, w):
    return [torch.zeros_like(a) for a in w]


def run_gpu(mvd, torch, dims, nchem, bc, w_host, plain, gamma=1.4):
    nx, ny, nz = dims
    ctx = mvd.Context(mvd.make_params(nx, ny, nz, nchem, bc, gamma))
    w = to_dev(torch, w_host)
    wd = zeros_dev(torch, w)
    rc = ctx.fe(w, wd, plain=plain, allow_illegal=True)
    out = [a.cpu().numpy() for a in wd]
    ctx.close()
    return rc, out


BCS = {
    "periodic": (0, 0, 0, 0, 0, 0),
    "neumann": (1, 1, 1, 1, 1, 1),
    "reflecting": (3, 3, 3, 3, 3, 3),
    "mixed": (0, 0, 1, 3, 3, 1),
}


@pytest.mark.parametrize("bcname", list(BCS))
@pytest.mark.parametrize("dims,nchem", [((12, 9, 7), 0), ((8, 6, 10), 2), ((7, 8, 6), 10),
                                        ((3, 5, 4), 0), ((37, 33, 9), 0), ((9, 7, 11), 1),
                                        ((17, 12, 5), 4), ((6, 9, 8), 6), ((70, 5, 4), 10),
                                        ((40, 33, 35), 10), ((33, 31, 17), 4)])
def test_fe_plain_bitwise_and_fast(mvd, torch_gpu, bcname, dims, nchem):
    torch = torch_gpu
    bc = BCS[bcname]
    print(f"[fe parity] dims={dims} nchem={nchem} bc={bcname}", flush=True)
    w = problems.smooth_state(*dims, nchem=nchem, seed=11, noise=0.02)
    g = orc.make_grid(*dims, nchem, bc)
    flag, ref = orc.fe(g, w)
    assert flag == 0
    rc, plain = run_gpu(mvd, torch, dims, nchem, bc, w, plain=True)
    assert rc == 0
    for v in range(5 + (1 if nchem else 0)):
        assert np.array_equal(plain[v].view(np.uint64), ref[v].view(np.uint64)), \
            f"plain kernel differs from oracle in field {v}"
    rc, fast = run_gpu(mvd, torch, dims, nchem, bc, w, plain=False)
    assert rc == 0
    for v in range(5):
        assert normwise_err(fast[v], ref[v]) <= TOL, f"field {v}"
    if nchem:
        r = ref[5].reshape(-1, nchem)
        f = fast[5].reshape(-1, nchem)
        for s in range(nchem):
            assert normwise_err(f[:, s], r[:, s]) <= TOL, f"species {s}"


def test_fe_blast_64(mvd, torch_gpu):
    """The 3D blast state (periodic, nchem=0) at 64^3 against the oracle."""
    torch = torch_gpu
    n = 64
    print(f"[fe blast] n={n}", flush=True)
    w = problems.blast3d(n, seed=1234) + [np.zeros(1)]
    g = orc.make_grid(n, n, n, 0, (0,) * 6)
    flag, ref = orc.fe(g, w)
    rc, fast = run_gpu(mvd, torch, (n, n, n), 0, (0,) * 6, w, plain=False)
    assert flag == 0 and rc == 0
    for v in range(5):
        assert normwise_err(fast[v], ref[v]) <= TOL


def test_fe_sod_strip(mvd, torch_gpu):
    """Sod shock tube layout 128x3x3 with Neumann boundaries (discontinuous data)."""
    torch = torch_gpu
    w = problems.sod(128) + [np.zeros(1)]
    bc = (1,) * 6
    g = orc.make_grid(128, 3, 3, 0, bc)
    flag, ref = orc.fe(g, w)
    rc, plain = run_gpu(mvd, torch, (128, 3, 3), 0, bc, w, plain=True)
    for v in range(5):
        assert np.array_equal(plain[v], ref[v])
    rc, fast = run_gpu(mvd, torch, (128, 3, 3), 0, bc, w, plain=False)
    for v in range(5):
        assert normwise_err(fast[v], ref[v]) <= TOL


@pytest.mark.parametrize("bad,bits", [("rho", 1), ("et", 2 | 4), ("p", 4)])
def test_fe_illegal_state(mvd, torch_gpu, bad, bits):
    """legal_state (euler3D.hpp:1405-1414): non-positive rho/et/p makes fe return -1."""
    torch = torch_gpu
    dims = (6, 5, 4)
    w = problems.smooth_state(*dims, seed=3)
    c = 37
    if bad == "rho":
        w[0][c] = -1.0
    elif bad == "et":
        w[4][c] = -1.0
    else:
        w[4][c] = 0.5 * (w[1][c] ** 2 + w[2][c] ** 2 + w[3][c] ** 2) / w[0][c] * 0.5
    g = orc.make_grid(*dims, 0)
    flag, _ = orc.fe(g, w)
    assert flag == -1
    for plain in (True, False):
        ctx = mvd.Context(mvd.make_params(*dims, 0))
        wd = to_dev(torch, w)
        out = zeros_dev(torch, wd)
        rc = ctx.fe(wd, out, plain=plain, allow_illegal=True)
        assert rc == mvd.MVD_ILLEGAL_STATE
        assert ctx.last_state_flags() & bits
        ctx.close()


def test_stability(mvd, torch_gpu):
    torch = torch_gpu
    dims = (20, 16, 12)
    w = problems.smooth_state(*dims, seed=5, noise=0.05)
    g = orc.make_grid(*dims, 0)
    ctx = mvd.Context(mvd.make_params(*dims, 0, cfl=0.3))
    wd = to_dev(torch, w)
    a = ctx.max_wave_speed(wd)
    assert a == orc.max_wave_speed(g, w)
    assert ctx.stability(wd) == orc.stability_dt(g, w, 0.3)
    ctx.close()


def test_fe_256_properties(mvd, torch_gpu):
    """At the benchmark size (256^3), size-independent properties: a constant state gives
    fe == 0 exactly; with periodic boundaries the domain sums of fe telescope to ~0;
    fast and plain kernels agree to 1e-12."""
    torch = torch_gpu
    n = 256
    print("[fe 256] properties", flush=True)
    ctx = mvd.Context(mvd.make_params(n, n, n, 0))
    N = n ** 3
    const = [torch.full((N,), v, dtype=torch.float64, device="cuda")
             for v in (1.3, 0.2, -0.1, 0.05, 4.0)] + [None]
    out = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(5)] + [None]
    ctx.fe(const, out)
    for v in range(5):
        assert torch.count_nonzero(out[v]).item() == 0
    w = [torch.from_numpy(a).to("cuda") for a in problems.blast3d(n, seed=99)] + [None]
    ctx.fe(w, out)
    fast = [o.clone() for o in out[:5]]
    ctx.fe(w, out, plain=True)
    for v in range(5):
        err = (fast[v] - out[v]).abs().max().item() / out[v].abs().max().item()
        assert err <= TOL
        s = fast[v].sum().item()
        assert abs(s) <= 1e-9 * fast[v].abs().sum().item()
    ctx.close()

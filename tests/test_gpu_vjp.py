"""Vector-Jacobian products on the GPU through the C ABI against the CPU oracle.

Chemistry: mvd_chem_vjp against orc_chem_vjp (itself pinned by central differences, the
conservation cotangents and the reference Jacobian, tests/test_oracle_vjp.py).
Tolerance 1e-10 per component relative to the largest magnitude of that component over
the batch: the GPU accumulates the reverse sweep in another order than the oracle's
forward-mode Jacobian.
"""
import importlib.util
import os

import numpy as np
import pytest

import orc
from conftest import PKG, normwise_err

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

pytestmark = pytest.mark.gpu
TOL = 1e-10


def chem_pair(mvd, torch, n, seed):
    y, _ = problems.chem_cells(n, seed=seed)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    rng = np.random.default_rng(seed)
    yo *= 1.0 + 0.05 * rng.standard_normal(yo.size)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    yd.copy_(torch.from_numpy(yo))
    # converge both temperature warm starts with the same RHS calls
    rd = torch.empty_like(yd)
    for _ in range(3):
        oc.rhs(yo)
        ctx.chem_rhs(yd, rd)
    return oc, yo, ctx, yd


@pytest.mark.parametrize("n,seed", [(4096, 11), (1000, 12)])
def test_chem_vjp_matches_oracle(mvd, torch_gpu, n, seed):
    torch = torch_gpu
    print(f"[chem vjp] n={n} seed={seed}", flush=True)
    oc, yo, ctx, yd = chem_pair(mvd, torch, n, seed)
    rng = np.random.default_rng(seed + 1)
    v = rng.standard_normal(n * 10)
    ref = oc.vjp(yo, v).reshape(n, 10)
    vd = torch.from_numpy(v).to("cuda")
    yb = torch.empty_like(yd)
    f = torch.empty_like(yd)
    ctx.chem_vjp(yd, vd, yb, f)
    got = yb.cpu().numpy().reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= TOL, s
    # the RHS returned alongside is mvd_chem_rhs's (the second call starts from the
    # temperature the first one stored, so agreement is to the last bits of T)
    f2 = torch.empty_like(yd)
    ctx.chem_rhs(yd, f2)
    a, b = f.cpu().numpy().reshape(n, 10), f2.cpu().numpy().reshape(n, 10)
    for s in range(10):
        assert normwise_err(a[:, s], b[:, s]) <= 1e-13, s


def test_chem_vjp_conservation_cotangents(mvd, torch_gpu):
    torch = torch_gpu
    n = 2048
    oc, yo, ctx, yd = chem_pair(mvd, torch, n, 13)
    sc = ctx.chem_state("scale", torch).cpu().numpy().reshape(n, 10)
    rng = np.random.default_rng(3)
    vr = torch.from_numpy(rng.standard_normal(n * 10)).to("cuda")
    yb = torch.empty_like(yd)
    ctx.chem_vjp(yd, vr, yb)
    ref = np.abs(yb.cpu().numpy()).reshape(n, 10).max(axis=1)
    vrm = np.abs(vr.cpu().numpy()).reshape(n, 10).max(axis=1)
    for c in ([2, 2, 1, 1, 1, 0, 0, 0, 0, 0.0], [0, 0, 0, 0, 0, 1, 1, 1, 0, 0.0],
              [0, 1, 0, 1, -1, 0, 1, 2, -1, 0.0]):
        v = (np.asarray(c)[None, :] * sc).reshape(-1)
        ctx.chem_vjp(yd, torch.from_numpy(v).to("cuda"), yb)
        got = np.abs(yb.cpu().numpy()).reshape(n, 10).max(axis=1)
        mag = np.abs(v).reshape(n, 10).max(axis=1) / vrm
        assert np.all(got <= 1e-11 * ref * mag + 1e-300)

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


# ---- fe() ----------------------------------------------------------------------------

@pytest.mark.parametrize("dims,nchem,bc", [((12, 10, 9), 0, (0,) * 6),
                                           ((11, 9, 8), 0, (1, 3, 3, 1, 3, 1)),
                                           ((10, 9, 8), 2, (0, 0, 1, 3, 0, 0)),
                                           ((9, 8, 10), 10, (3, 1, 0, 0, 1, 3)),
                                           # (no Dirichlet: its negated ghost density
                                           # makes the reference's Roe average divide by
                                           # sqrt(rho) = 0, so fe() is not finite there)
                                           ((7, 9, 6), 1, (3, 3, 0, 0, 1, 1)),
                                           ((8, 7, 9), 4, (0, 0, 3, 3, 1, 1)),
                                           ((6, 8, 7), 6, (1, 1, 3, 3, 0, 0)),
                                           ((70, 5, 4), 10, (0,) * 6),
                                           # several segments per y and z column and
                                           # several x-lines per wave
                                           ((18, 24, 16), 1, (1, 3, 0, 0, 3, 1)),
                                           # x-lines of five cells: two periodic ghost
                                           # points share a source cell (the x fix-up's
                                           # one-pair-at-a-time path)
                                           ((5, 9, 7), 2, (0,) * 6)])
def test_fe_vjp_matches_oracle(mvd, torch_gpu, dims, nchem, bc):
    torch = torch_gpu
    print(f"[fe vjp] dims={dims} nchem={nchem} bc={bc}", flush=True)
    nx, ny, nz = dims
    w = problems.smooth_state(nx, ny, nz, nchem, seed=3)
    rng = np.random.default_rng(4)
    ncell = nx * ny * nz
    v = [rng.standard_normal(ncell) for _ in range(5)] + \
        [rng.standard_normal(max(nchem, 1) * ncell)]
    g = orc.make_grid(nx, ny, nz, nchem, bc)
    ref = orc.fe_vjp(g, w, v)
    ctx = mvd.Context(mvd.make_params(nx, ny, nz, nchem, bc))
    wd = [torch.from_numpy(a).to("cuda") for a in w]
    vd = [torch.from_numpy(a).to("cuda") for a in v]
    if nchem == 0:
        wd[5] = vd[5] = None
    wb = [torch.empty(ncell, dtype=torch.float64, device="cuda") for _ in range(5)] + \
         [torch.empty(nchem * ncell, dtype=torch.float64, device="cuda") if nchem else None]
    ctx.fe_vjp(wd, vd, wb)
    for q in range(5):
        assert normwise_err(wb[q].cpu().numpy(), ref[q]) <= TOL, q
    if nchem:
        got = wb[5].cpu().numpy().reshape(ncell, nchem)
        r = ref[5].reshape(ncell, nchem)
        for s in range(nchem):
            assert normwise_err(got[:, s], r[:, s]) <= TOL, ("chem", s)


def test_fe_vjp_dot_product_64(mvd, torch_gpu):
    """<v, (fe(w+hu) - fe(w-hu))/2h> = <wbar, u> with the fast fe() kernels at 64^3."""
    torch = torch_gpu
    n = 64
    print(f"[fe vjp dot] n={n}", flush=True)
    w = [torch.from_numpy(a).to("cuda") for a in problems.blast3d(n, seed=5)[:5]] + [None]
    gen = torch.Generator(device="cuda").manual_seed(7)
    u = [torch.randn(n ** 3, dtype=torch.float64, device="cuda", generator=gen) * a.abs().max()
         for a in w[:5]] + [None]
    v = [torch.randn(n ** 3, dtype=torch.float64, device="cuda", generator=gen)
         for _ in range(5)] + [None]
    ctx = mvd.Context(mvd.make_params(n, n, n, 0, (0,) * 6))
    wb = [torch.empty_like(a) for a in w[:5]] + [None]
    ctx.fe_vjp(w, v, wb)
    h = 1e-6
    fp = [torch.empty_like(a) for a in w[:5]] + [None]
    fm = [torch.empty_like(a) for a in w[:5]] + [None]
    ctx.fe([a + h * b for a, b in zip(w[:5], u[:5])] + [None], fp)
    ctx.fe([a - h * b for a, b in zip(w[:5], u[:5])] + [None], fm)
    lhs = sum(float(torch.dot(vv, (a - b) / (2 * h))) for vv, a, b in zip(v[:5], fp[:5], fm[:5]))
    rhs = sum(float(torch.dot(a, b)) for a, b in zip(wb[:5], u[:5]))
    mag = sum(float(torch.dot(a.abs(), b.abs())) for a, b in zip(wb[:5], u[:5]))
    assert abs(lhs - rhs) <= 1e-6 * mag, (lhs, rhs, mag)

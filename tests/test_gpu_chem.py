"""Primordial chemistry on the GPU through the C ABI against the CPU oracle.

Tolerance: 1e-12, per species, relative to the largest magnitude of that quantity over the
batch (device log/exp/pow differ from glibc by an ulp; the expressions are otherwise
evaluated in the reference's order without FMA contraction).
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
TOL = 1e-12


def gpu_chem(mvd, torch, n, y):
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    return ctx, yd


@pytest.mark.parametrize("n,seed", [(4096, 1), (1000, 2)])
def test_chem_prepare_rhs_jac(mvd, torch_gpu, n, seed):
    torch = torch_gpu
    print(f"[chem parity] n={n} seed={seed}", flush=True)
    y, _ = problems.chem_cells(n, seed=seed)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    ctx, yd = gpu_chem(mvd, torch, n, y)
    assert np.array_equal(yd.cpu().numpy(), yo)              # chem <- 1
    for name in ("scale", "inv_scale", "mdensity", "inv_mdensity", "cie_oda", "h2_oda"):
        got = ctx.chem_state(name, torch).cpu().numpy()
        assert normwise_err(got, oc.array(name)) <= TOL, name
    # two RHS calls: the second is warm-started from the first call's temperature
    for call in range(2):
        rd = torch.empty_like(yd)
        ctx.chem_rhs(yd, rd)
        ref = oc.rhs(yo)
        r = rd.cpu().numpy().reshape(n, 10)
        rr = ref.reshape(n, 10)
        for s in range(10):
            assert normwise_err(r[:, s], rr[:, s]) <= TOL, (call, s)
        Ts = ctx.chem_state("Ts", torch).cpu().numpy()
        assert np.max(np.abs(Ts - oc.array("Ts")) / oc.array("Ts")) <= TOL
        dT = ctx.chem_state("dTs_ge", torch).cpu().numpy()
        assert normwise_err(dT, oc.array("dTs_ge")) <= TOL
    Jd = torch.empty(100 * n, dtype=torch.float64, device="cuda")
    ctx.chem_jac(yd, Jd)
    J = Jd.cpu().numpy().reshape(n, 10, 10).transpose(0, 2, 1)
    Jr = oc.jac(yo)
    for a in range(10):
        for b in range(10):
            assert normwise_err(J[:, a, b], Jr[:, a, b]) <= TOL, (a, b)
    ctx.close()


def test_temperature_early_exit_on_device(mvd, torch_gpu):
    """The device temperature solve ends its ten Newton steps at the first bitwise repeat of
    the iteration (period <= 4); the self-check entry point runs that form and the full ten
    steps side by side on the device and must find no cell whose (T, dTs_ge) differ, from
    cold, true, converged and perturbed warm starts. The oracle's restatement of the exit
    shows that these inputs hold repeats of every period up to 4."""
    torch = torch_gpu
    n = 4096
    print(f"[temperature exit] n={n}", flush=True)
    y, T_true = problems.chem_cells(n, seed=21)
    ctx, yd = gpu_chem(mvd, torch, n, y)
    rd = torch.empty_like(yd)
    ctx.chem_rhs(yd, rd)
    Tw = ctx.chem_state("Ts", torch).cpu().numpy()
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    periods = np.zeros(5, dtype=np.int64)
    for Ts0 in (np.full(n, 1000.0), T_true, Tw, Tw * (1 + 1e-12), Tw * (1 - 1e-9), Tw * 1.01):
        T0 = torch.from_numpy(np.ascontiguousarray(Ts0)).to("cuda")
        assert ctx.temperature_check(yd, T0) == 0
        _, _, per = oc.temperature_exit_check(yo, np.ascontiguousarray(Ts0), 4)
        periods += np.bincount(per, minlength=5)
    assert np.all(periods[1:] > 0), periods
    ctx.close()


def test_chem_scaling_roundtrip(mvd, torch_gpu):
    torch = torch_gpu
    n = 512
    y, _ = problems.chem_cells(n, seed=4)
    ctx, yd = gpu_chem(mvd, torch, n, y)
    z = torch.from_numpy(np.random.default_rng(1).uniform(0.5, 1.5, 10 * n)).to("cuda")
    z0 = z.clone()
    ctx.chem_apply_scaling(z)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    zo = z0.cpu().numpy().copy()
    oc_apply = zo * oc.array("scale")
    assert np.array_equal(z.cpu().numpy(), oc_apply)
    ctx.chem_unapply_scaling(z)
    assert np.array_equal(z.cpu().numpy(), oc_apply / oc.array("scale"))
    ctx.close()


def test_fslow_matches_oracle(mvd, torch_gpu):
    """fslow: scaled chem, et rebuilt from the gas energy, fe() with 10 species, gas-energy
    slot takes etdot, etdot zeroed, chem unscaled (multirate_chem_hydro_main.cpp:996-1083)."""
    torch = torch_gpu
    dims = (12, 10, 8)
    n = 12 * 10 * 8
    print("[fslow parity]", flush=True)
    w = problems.smooth_state(*dims, nchem=10, seed=6, noise=0.01)
    y, _ = problems.chem_cells(n, seed=7)
    # gas energy per unit mass consistent with the fluid state (code units, EnergyUnits=1)
    rho, mx, my, mz, et = w[:5]
    ge = (et - 0.5 * (mx ** 2 + my ** 2 + mz ** 2) / rho) / rho
    y = y.reshape(n, 10)
    y[:, 9] = ge * rho
    y = y.reshape(-1).copy()
    bc = (3,) * 6
    g = orc.make_grid(*dims, 10, bc)
    oc = orc.Chem(n)
    chem_o = y.copy()
    oc.prepare(chem_o)
    wo = [a.copy() for a in w[:5]] + [chem_o]
    flag, ref = orc.fslow(g, oc, 1.0, wo)
    assert flag == 0
    ctx = mvd.Context(mvd.make_params(*dims, 10, bc))
    ctx.chem_setup(orc.load_tables(), n)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in w[:5]] + \
        [torch.from_numpy(y.copy()).to("cuda")]
    ctx.chem_prepare(wd[5])
    out = [torch.empty_like(a) for a in wd]
    ctx.fslow(wd, out)
    for v in range(4):
        assert normwise_err(out[v].cpu().numpy(), ref[v]) <= TOL, v
    assert torch.count_nonzero(out[4]).item() == 0
    got = out[5].cpu().numpy().reshape(n, 10)
    rr = ref[5].reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], rr[:, s]) <= TOL, s
    # side effects on w: et overwritten, chem back to (scale*x)/scale
    assert normwise_err(wd[4].cpu().numpy(), wo[4]) <= 1e-15
    assert np.array_equal(wd[5].cpu().numpy(), wo[5])
    ctx.close()


def test_chem_block_ls_matches_scipy(mvd, torch_gpu):
    """mvd_chem_lsetup / mvd_chem_lsolve (A = I - gamma J per cell, partial pivoting) against
    scipy's getrf/getrs on the same blocks, on the Jacobian of config-3 cells at the
    SDIRK3 diagonal of a 1e3 s step; then a block made exactly singular is reported."""
    import scipy.linalg as sl
    torch = torch_gpu
    n = 2048
    print(f"[chem ls] n={n}", flush=True)
    y, _ = problems.chem_cells(n, seed=21)
    ctx, yd = gpu_chem(mvd, torch, n, y)
    rd = torch.empty_like(yd)
    ctx.chem_rhs(yd, rd)
    Jd = torch.empty(100 * n, dtype=torch.float64, device="cuda")
    ctx.chem_jac(yd, Jd)
    gamma = 1e3 * 0.435866521508459
    assert ctx.chem_lsetup(gamma, Jd) == 0
    rng = np.random.default_rng(5)
    b = rng.standard_normal(10 * n)
    bd = torch.from_numpy(b).to("cuda")
    xd = torch.empty_like(bd)
    ctx.chem_lsolve(bd, xd)
    x = xd.cpu().numpy().reshape(n, 10)
    Jc = Jd.cpu().numpy().reshape(n, 10, 10).transpose(0, 2, 1)      # [cell][row][col]
    A = np.eye(10)[None] - gamma * Jc
    bb = b.reshape(n, 10)
    xr = np.stack([sl.lu_solve(sl.lu_factor(A[i]), bb[i]) for i in range(n)])
    res = np.abs(np.einsum("nij,nj->ni", A, x) - bb).max(axis=1)
    scale = (np.abs(A) * np.abs(x)[:, None, :]).sum(axis=2).max(axis=1)
    assert np.all(res <= 1e-13 * scale)
    rel = np.abs(x - xr).max(axis=1) / np.abs(xr).max(axis=1)
    cond = np.linalg.cond(A)
    assert np.all(rel <= 1e-15 * cond + 1e-13)
    # in place (x aliases b)
    ctx.chem_lsolve(bd, bd)
    assert np.array_equal(bd.cpu().numpy().reshape(n, 10), x)
    # two singular blocks: J = I / gamma makes A = 0
    Jm = Jd.clone().reshape(n, 10, 10)
    Jm[3] = torch.eye(10, dtype=torch.float64, device="cuda") / gamma
    Jm[17] = torch.eye(10, dtype=torch.float64, device="cuda") / gamma
    assert ctx.chem_lsetup(gamma, Jm.reshape(-1)) == 2
    with pytest.raises(mvd.MvdError):
        ctx.chem_lsolve(bd, xd)
    ctx.close()

"""Pins of the oracle's chemistry vector-Jacobian product (orc_chem_vjp):

* against central finite differences of the oracle RHS (h = 1e-7 relative; every
  evaluation re-solves the temperature from the converged warm start), which pins the
  whole derivative including the temperature path;
* the three linear invariants: the cotangent c o scale (c = H-nuclei, He-nuclei or charge
  weights) must give a zero product for any state, because c . f_phys = 0 identically;
* at fixed rates (temperature path dropped) the product equals J_ref^T v on the entries
  where the reference's analytic Jacobian is the exact derivative: species rows x species
  columns (calculate_jacobian_cvklu, raja_primordial_network.cpp:1518-1867). Its
  gas-energy row is not: it differs from the derivative of the energy RHS by up to ~5e-6
  relative (the derivative here is the one the finite differences confirm), so the
  cotangents of that check have no energy component.
"""
import importlib.util
import os

import numpy as np
import pytest

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)


def prepared(n, seed, **kw):
    y, _ = problems.chem_cells(n, seed=seed, **kw)
    ch = orc.Chem(n)
    ys = y.copy()
    ch.prepare(ys)
    ch.extra()
    # perturb away from the all-ones scaled state so every entry is exercised
    rng = np.random.default_rng(seed)
    ys *= 1.0 + 0.05 * rng.standard_normal(ys.size)
    ch.rhs(ys)                                   # converge the temperature warm start
    for _ in range(2):
        ch.rhs(ys)
    return ch, ys


def fd_jacobian(ch, ys, rel=1e-7):
    n = ch.n
    Ts0 = ch.array("Ts").copy()
    J = np.zeros((n, 10, 10))
    for q in range(10):
        h = rel * np.maximum(np.abs(ys.reshape(n, 10)[:, q]), 1e-30)
        yp = ys.copy().reshape(n, 10)
        ym = ys.copy().reshape(n, 10)
        yp[:, q] += h
        ym[:, q] -= h
        ch.array("Ts")[:] = Ts0
        fp = ch.rhs(yp.reshape(-1)).reshape(n, 10)
        ch.array("Ts")[:] = Ts0
        fm = ch.rhs(ym.reshape(-1)).reshape(n, 10)
        J[:, :, q] = (fp - fm) / (2 * h[:, None])
    ch.array("Ts")[:] = Ts0
    return J


@pytest.mark.parametrize("seed,kw", [(3, {}), (4, {"temp": (2e3, 9e3)}),
                                     (5, {"ntot": (1e1, 1e2), "temp": (10.0, 300.0)})])
def test_vjp_matches_central_differences(seed, kw):
    n = 64
    ch, ys = prepared(n, seed, **kw)
    J = fd_jacobian(ch, ys)
    rng = np.random.default_rng(seed + 100)
    for _ in range(3):
        v = rng.standard_normal(n * 10)
        yb = ch.vjp(ys, v).reshape(n, 10)
        ref = np.einsum("nsq,ns->nq", J, v.reshape(n, 10))
        scale = np.einsum("nsq,ns->nq", np.abs(J), np.abs(v.reshape(n, 10))).max(axis=1) + 1e-300
        err = np.abs(yb - ref).max(axis=1) / scale
        # FD truncation and rounding at h = 1e-7: ~1e-6; a cell whose temperature sits on a
        # table-bin edge (kink of the interpolant) may exceed it
        assert np.mean(err < 1e-5) > 0.95 and np.median(err) < 1e-7, (err.max(), np.median(err))


def test_vjp_conservation_cotangents_give_zero():
    n = 256
    ch, ys = prepared(n, 7)
    sc = ch.array("scale").reshape(n, 10)
    weights = [np.array([2, 2, 1, 1, 1, 0, 0, 0, 0, 0.0]),           # H nuclei
               np.array([0, 0, 0, 0, 0, 1, 1, 1, 0, 0.0]),           # He nuclei
               np.array([0, 1, 0, 1, -1, 0, 1, 2, -1, 0.0])]         # charge
    rng = np.random.default_rng(1)
    vr = rng.standard_normal(n * 10)
    ref = np.abs(ch.vjp(ys, vr)).reshape(n, 10).max(axis=1)
    for c in weights:
        v = (c[None, :] * sc).reshape(-1)
        yb = ch.vjp(ys, v).reshape(n, 10)
        # relative to the product of a unit random cotangent of the same scale
        mag = np.abs(v).reshape(n, 10).max(axis=1) / np.abs(vr).reshape(n, 10).max(axis=1)
        assert np.all(np.abs(yb).max(axis=1) <= 1e-11 * ref * mag + 1e-300)


def test_fixed_rate_vjp_equals_reference_jacobian_on_species_columns():
    n = 64
    ch, ys = prepared(n, 9)
    rng = np.random.default_rng(2)
    v = rng.standard_normal((n, 10))
    v[:, 9] = 0.0                        # species rows only
    v = v.reshape(-1)
    J = ch.jac(ys)                       # [cell, row, col], from the rates of the last RHS
    yb = ch.vjp(ys, v, fixed_T=True).reshape(n, 10)
    ref = np.einsum("nsq,ns->nq", J, v.reshape(n, 10))
    scale = np.einsum("nsq,ns->nq", np.abs(J), np.abs(v.reshape(n, 10))).max(axis=1)[:, None]
    err = np.abs(yb[:, :9] - ref[:, :9]) / scale
    assert err.max() < 1e-12, err.max()

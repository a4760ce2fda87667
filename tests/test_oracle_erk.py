"""Explicit Runge-Kutta evolution on the CPU oracle (configuration 1 path), pinned by:
the Bogacki-Shampine table's order conditions (exact rationals) and the library's copy of
it; third-order convergence in time on a smooth wave; and the Sod problem at t = 0.2
against the exact Riemann solution (Toro, Riemann Solvers and Numerical Methods for Fluid
Dynamics, ch. 4) -- plateau values and the L1 density error.
"""
import importlib.util
import os

import numpy as np

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)


def test_bs3_order_conditions_and_library_table(mvd):
    A, b, c = orc.bs3_table()
    assert np.allclose(A.sum(axis=1), c, atol=1e-16, rtol=0)
    assert abs(b.sum() - 1) < 1e-15
    assert abs(b @ c - 1 / 2) < 1e-15
    assert abs(b @ c ** 2 - 1 / 3) < 1e-15
    assert abs(b @ A @ c - 1 / 6) < 1e-15
    assert np.array_equal(A[3, :3], b[:3]) and b[3] == 0        # first same as last
    t = mvd.bs3()
    assert t.stages == 4
    LA = np.array(t.A[:]).reshape(6, 6)[:4, :4]
    assert np.array_equal(LA, A)
    assert np.array_equal(np.array(t.b[:4]), b) and np.array_equal(np.array(t.c[:4]), c)


def smooth_wave(nx):
    """Periodic acoustic-like wave along x on nx x 3 x 3 (all fields smooth)."""
    x = (np.arange(nx) + 0.5) / nx
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x)
    u = 0.1 * np.cos(2 * np.pi * x)
    p = 1.0 + 0.1 * np.sin(2 * np.pi * x + 0.3)
    f = lambda a: np.ascontiguousarray(np.broadcast_to(a, (3, 3, nx))).reshape(-1)
    mx = rho * u
    et = p / 0.4 + 0.5 * rho * u * u
    return [f(rho), f(mx), f(np.zeros(nx)), f(np.zeros(nx)), f(et), np.zeros(1)]


def evolve(g, w0, h, nsteps):
    w = [a.copy() for a in w0]
    tab = orc.erk(*orc.bs3_table())
    assert orc.erk_evolve(g, tab, h, nsteps, w) == 0
    return w


def test_erk3_third_order_in_time():
    g = orc.make_grid(32, 3, 3, 0, (0,) * 6)
    w0 = smooth_wave(32)
    T = 0.08
    ref = evolve(g, w0, T / 128, 128)
    errs = []
    for k in (8, 16, 32):
        w = evolve(g, w0, T / k, k)
        errs.append(max(np.abs(a - r).max() for a, r in zip(w[:5], ref[:5])))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(2.7 < r < 3.3 for r in rates), (errs, rates)


def sod_exact(x, t, gamma=1.4, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), x0=0.5):
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    g1, g2 = (gamma - 1) / (2 * gamma), (gamma + 1) / (2 * gamma)

    def fk(p, rk, pk, ck):
        if p > pk:       # shock
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            q = np.sqrt(A / (p + B))
            return (p - pk) * q, q * (1 - 0.5 * (p - pk) / (p + B))
        f = 2 * ck / (gamma - 1) * ((p / pk) ** g1 - 1)   # rarefaction
        return f, 1 / (rk * ck) * (p / pk) ** (-g2)

    p = 0.5 * (pl + pr)
    for _ in range(100):
        f1, d1 = fk(p, rl, pl, cl)
        f2, d2 = fk(p, rr, pr, cr)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p -= dp
        if abs(dp) < 1e-15 * p:
            break
    u = 0.5 * (ul + ur) + 0.5 * (fk(p, rr, pr, cr)[0] - fk(p, rl, pl, cl)[0])
    rsl = rl * (p / pl) ** (1 / gamma)                     # left rarefaction
    rsr = rr * ((p / pr + (gamma - 1) / (gamma + 1)) / ((gamma - 1) / (gamma + 1) * p / pr + 1))
    s_shock = ur + cr * np.sqrt(g2 * p / pr + g1)
    csl = cl * (p / pl) ** g1
    s = (x - x0) / t
    rho = np.empty_like(x)
    for i, si in enumerate(s):
        if si < ul - cl:
            rho[i] = rl
        elif si < u - csl:                                  # inside the fan
            uf = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * ul + si)
            cf = uf - si
            rho[i] = rl * (cf / cl) ** (2 / (gamma - 1))
        elif si < u:
            rho[i] = rsl
        elif si < s_shock:
            rho[i] = rsr
        else:
            rho[i] = rr
    return rho, p, u, rsl, rsr


def test_sod_oracle_vs_exact_riemann():
    nx = 128
    g = orc.make_grid(nx, 3, 3, 0, (1,) * 6)
    w0 = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in problems.sod(nx)[:5]]
    w = evolve(g, w0 + [np.zeros(1)], 1e-4, 2000)
    rho = w[0].reshape(3, 3, nx)[1, 1]
    mx = w[1].reshape(3, 3, nx)[1, 1]
    et = w[4].reshape(3, 3, nx)[1, 1]
    p = 0.4 * (et - 0.5 * mx * mx / rho)
    x = (np.arange(nx) + 0.5) / nx
    rex, pstar, ustar, rsl, rsr = sod_exact(x, 0.2)
    assert abs(pstar - 0.30313) < 1e-4 and abs(ustar - 0.92745) < 1e-4
    assert np.mean(np.abs(rho - rex)) < 0.01
    # plateaus, away from the smeared waves: left star state between the fan tail
    # (x = 0.486) and the contact (x = 0.686), right star state between the contact and
    # the shock (x = 0.850)
    assert np.all(np.abs(rho[(x > 0.52) & (x < 0.63)] - rsl) < 2e-3)
    assert np.all(np.abs(rho[(x > 0.72) & (x < 0.82)] - rsr) < 2e-3)
    assert np.all(np.abs(p[(x > 0.52) & (x < 0.82)] - pstar) < 2e-3)
    # the y and z copies of the line stay identical (Neumann in y, z)
    r3 = w[0].reshape(3, 3, nx)
    assert np.array_equal(r3[0, 0], r3[2, 2]) and np.array_equal(r3[0, 1], r3[1, 1])

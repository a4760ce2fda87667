"""Discrete adjoint of the chemistry DIRK steps on the CPU oracle (orc_chem_dirk_adjoint),
pinned by:
  * the three linear invariants of the species RHS (H nuclei, He nuclei, charge): every
    Runge-Kutta step preserves them, so their cotangents pass through unchanged
    (lam0 = lamT to round-off) whatever the trajectory;
  * per-cell central differences of y_N . lamT (relative step 1e-7) with Newton converged to
    round-off, so that the computed map is the idealised one the adjoint differentiates;
  * a dot-product test along a random direction;
  * the exact Jacobian it uses against central differences of the RHS.
"""
import importlib.util
import os

import numpy as np

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

N, STEPS, H = 48, 6, 200.0
INVARIANTS = ([2, 2, 1, 1, 1, 0, 0, 0, 0, 0.0], [0, 0, 0, 0, 0, 1, 1, 1, 0, 0.0],
              [0, 1, 0, 1, -1, 0, 1, 2, -1, 0.0])


def setup(seed=31):
    y, _ = problems.chem_cells(N, seed=seed)
    ch = orc.Chem(N)
    ys = y.copy()
    ch.prepare(ys)
    ch.extra()
    rng = np.random.default_rng(seed)
    ys *= 1.0 + 0.3 * rng.uniform(-1, 1, ys.size)     # away from the prepared y = 1
    Ts0 = ch.array("Ts").copy()
    return ch, ys, Ts0


def tight_opts():
    return orc.dirk_opts(rtol=1e-13, atol=1e-30, newton_tol=0.1, max_iters=40, max_setups=6)


def run_forward(ch, Ts0, y0, opts):
    ch.array("Ts")[:] = Ts0
    y = y0.copy()
    for _ in range(STEPS):
        failed, _ = orc.dirk_step(ch, opts, H, y)
        assert failed == 0
    return y


def test_invariant_cotangents_pass_through():
    ch, y0, Ts0 = setup()
    sc = ch.array("scale").reshape(N, 10)
    for c in INVARIANTS:
        lamT = (np.asarray(c)[None, :] * sc).reshape(-1)
        y = y0.copy()
        ch.array("Ts")[:] = Ts0
        failed, lam0 = orc.dirk_adjoint(ch, orc.dirk_opts(), H, STEPS, y, lamT)
        assert failed == 0
        err = np.abs(lam0 - lamT).reshape(N, 10).max(axis=1)
        assert np.all(err <= 1e-12 * np.abs(lamT).reshape(N, 10).max(axis=1))


def test_adjoint_matches_central_differences():
    ch, y0, Ts0 = setup(32)
    opts = tight_opts()
    rng = np.random.default_rng(7)
    lamT = rng.standard_normal(10 * N)
    y = y0.copy()
    ch.array("Ts")[:] = Ts0
    failed, lam0 = orc.dirk_adjoint(ch, opts, H, STEPS, y, lamT)
    assert failed == 0
    assert np.array_equal(y, run_forward(ch, Ts0, y0, opts))   # forward part = plain steps
    L0 = lam0.reshape(N, 10)
    for j in range(10):
        d = 1e-7 * np.abs(y0.reshape(N, 10)[:, j])
        e = np.zeros((N, 10))
        e[:, j] = d
        yp = run_forward(ch, Ts0, y0 + e.reshape(-1), opts)
        ym = run_forward(ch, Ts0, y0 - e.reshape(-1), opts)
        fd = ((yp - ym).reshape(N, 10) * lamT.reshape(N, 10)).sum(axis=1) / (2 * d)
        scale = np.abs(L0).max(axis=1)
        assert np.all(np.abs(fd - L0[:, j]) <= 2e-5 * scale + 1e-300), j


def test_adjoint_dot_product():
    ch, y0, Ts0 = setup(33)
    opts = tight_opts()
    rng = np.random.default_rng(8)
    lamT = rng.standard_normal(10 * N)
    u = rng.standard_normal(10 * N) * np.abs(y0)
    y = y0.copy()
    ch.array("Ts")[:] = Ts0
    _, lam0 = orc.dirk_adjoint(ch, opts, H, STEPS, y, lamT)
    eps = 1e-7
    yp = run_forward(ch, Ts0, y0 + eps * u, opts)
    ym = run_forward(ch, Ts0, y0 - eps * u, opts)
    lhs = lamT @ (yp - ym) / (2 * eps)
    rhs = lam0 @ u
    assert abs(lhs - rhs) <= 1e-5 * np.abs(lam0 * u).sum()


def test_exact_jacobian_matches_central_differences():
    ch, y0, Ts0 = setup(34)
    for i in (0, 7, 30):
        yc = y0.reshape(N, 10)[i].copy()
        J, T = orc.exact_jac_cell(ch, i, yc, Ts0[i])
        for q in range(10):
            d = 1e-6 * abs(yc[q])
            e = np.zeros(10)
            e[q] = d
            ch.array("Ts")[:] = Ts0
            yp = np.tile(y0.reshape(N, 10), (1, 1))
            ym = yp.copy()
            yp[i] += e
            ym[i] -= e
            fp = ch.rhs(yp.reshape(-1)).reshape(N, 10)[i]
            ch.array("Ts")[:] = Ts0
            fm = ch.rhs(ym.reshape(-1)).reshape(N, 10)[i]
            fd = (fp - fm) / (2 * d)
            assert np.all(np.abs(fd - J[:, q]) <= 1e-5 * np.abs(J).max(axis=1) + 1e-300), (i, q)

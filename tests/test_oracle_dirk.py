"""Pins of the oracle's per-cell SDIRK/Newton chemistry step (oracle/orc_dirk.cpp) and of
its 10x10 LU, by checks that do not restate the oracle:

* LU with partial pivoting against LAPACK getrf (scipy.linalg.lu_factor): same pivots,
  same factors to rounding, and the solve against numpy.linalg.solve;
* the SDIRK3 table against its order conditions and L-stability, and the library's table
  against the one derived here from the defining cubic;
* third-order convergence of the whole step machinery on a smooth chemistry case (a wrong
  stage coefficient, stage predictor or derivative drops the order);
* exact preservation of the three linear invariants (H nuclei, He nuclei, charge): every
  Newton update keeps w.z = w.r for weights w with w^T J = 0, so the step conserves them
  to rounding for any Newton tolerance;
* the failure path (Newton budget exhausted) reported per cell.

The reference's integrator, Newton and LU outputs come from SUNDIALS/MAGMA, which its own
tests do not pin; these checks are what fixes them here.
"""
import importlib.util
import os

import numpy as np
import pytest
import scipy.linalg

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

MH, M_AMU = problems.MH, problems.M_AMU


def _matrices(seed, count=40):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        M = rng.standard_normal((10, 10))
        if k % 4 == 1:
            M[0, 0] = 1e-12              # forces a pivot at the first step
        if k % 4 == 2:
            M = np.eye(10) - 1e3 * rng.standard_normal((10, 10)) * 1e-2   # I - hg J like
        if k % 4 == 3:
            M[:, 3] *= 1e-8              # badly scaled column
        out.append(M)
    return out


def test_lu_matches_lapack_getrf():
    for M in _matrices(11):
        b = np.arange(1.0, 11.0)
        LU, piv, x = orc.lu(M, b)
        lu_ref, piv_ref = scipy.linalg.lu_factor(M)
        assert np.array_equal(piv, piv_ref)
        assert np.max(np.abs(LU - lu_ref)) <= 1e-12 * np.max(np.abs(lu_ref))
        xr = np.linalg.solve(M, b)
        assert np.max(np.abs(x - xr)) <= 1e-9 * np.max(np.abs(xr))
        # P M = L U from the oracle's own factors
        L = np.tril(LU, -1) + np.eye(10)
        U = np.triu(LU)
        PM = M.copy()
        for k in range(10):
            PM[[k, piv[k]]] = PM[[piv[k], k]]
        assert np.max(np.abs(L @ U - PM)) <= 1e-13 * np.max(np.abs(M)) * 10


def test_sdirk3_order_conditions_and_L_stability():
    A, c = orc.sdirk3_table()
    A3 = A[:3, :3]
    b = A3[2]                                   # stiffly accurate: b = last row
    c3 = A3.sum(axis=1)
    assert np.allclose(c3, c[:3], rtol=0, atol=1e-15)
    assert abs(b.sum() - 1.0) < 1e-14
    assert abs(b @ c3 - 0.5) < 1e-14
    assert abs(b @ c3 ** 2 - 1.0 / 3.0) < 1e-14
    assert abs(b @ A3 @ c3 - 1.0 / 6.0) < 1e-14
    # R(z) = 1 + z b^T (I - z A)^{-1} 1 -> 0 as z -> -inf (L-stable), |R| <= 1 on z < 0
    for z in (-1e12, -1e8):
        R = 1 + z * b @ np.linalg.solve(np.eye(3) - z * A3, np.ones(3))
        assert abs(R) < 1e-6
    for z in -np.logspace(-3, 6, 50):
        R = 1 + z * b @ np.linalg.solve(np.eye(3) - z * A3, np.ones(3))
        assert abs(R) <= 1.0 + 1e-14


def test_library_sdirk3_table(mvd):
    o = mvd.sdirk3()
    A, c = orc.sdirk3_table()
    assert o.nstage == 3
    # np.roots (companion-matrix eigenvalues) is accurate to a few ulp of the coefficients
    LA = np.array(list(o.A)).reshape(7, 7)
    assert np.max(np.abs(LA[:4, :4] - A)) <= 1e-14 and not LA[4:].any() and not LA[:, 4:].any()
    assert np.max(np.abs(np.array(list(o.c))[:4] - c)) <= 1e-14


def _order_conditions(A):
    """Residuals of the eight fourth-order conditions of the stiffly accurate table A."""
    b = A[-1]
    c = A.sum(axis=1)
    Ac = A @ c
    return np.array([b.sum() - 1, b @ c - 1 / 2, b @ c ** 2 - 1 / 3, b @ Ac - 1 / 6,
                     b @ c ** 3 - 1 / 4, b @ (c * Ac) - 1 / 8, b @ (A @ c ** 2) - 1 / 12,
                     b @ (A @ Ac) - 1 / 24])


def test_ark437_table_order_conditions_and_L_stability():
    """ARK4(3)7L[2]SA's ESDIRK part: the published nodes are the row sums, all eight
    fourth-order conditions hold (six of them independent of the two that fixed b_5, b_6),
    the method is L-stable; a wrong digit in one entry breaks the conditions."""
    A, c = orc.ark437_table()
    assert np.allclose(A.sum(axis=1), c, rtol=0, atol=1e-15)
    assert np.all(np.diag(A)[1:] == 0.1235) and A[0, 0] == 0.0 and not np.triu(A, 1).any()
    assert np.max(np.abs(_order_conditions(A))) < 1e-15
    b = A[-1]
    for z in (-1e12, -1e8):
        R = 1 + z * b @ np.linalg.solve(np.eye(7) - z * A, np.ones(7))
        assert abs(R) < 1e-6
    for z in -np.logspace(-3, 6, 50):
        R = 1 + z * b @ np.linalg.solve(np.eye(7) - z * A, np.ones(7))
        assert abs(R) <= 1.0 + 1e-14
    bad = A.copy()
    bad[4, 2] *= 1 + 1e-9                       # one entry off in its tenth digit
    bad[4, 4] -= bad[4].sum() - A[4].sum()      # (keeping the node)
    assert np.max(np.abs(_order_conditions(bad))) > 1e-12


def test_library_ark437_table(mvd):
    o = mvd.ark437()
    A, c = orc.ark437_table()
    assert o.nstage == 7
    LA = np.array(list(o.A)).reshape(7, 7)
    assert np.array_equal(LA, A)                # the same doubles (correctly rounded)
    assert np.max(np.abs(np.array(list(o.c)) - c)) <= 1e-15


def test_oracle_esdirk_fourth_order():
    """Global error of N fixed ARK4(3)7L[2]SA-ESDIRK steps against N=512 steps: ratio ~ 16
    per halving (14.6 and 15.8 here, where SDIRK3 gives ~7), which checks the oracle's
    explicit first stage and stage sums; on the warm gas of the third-order test."""
    n, T = 32, 1e8
    y0, _ = problems.chem_cells(n, seed=7, ntot=(1e-1, 1e1), temp=(3e3, 8e3))
    A, c = orc.ark437_table()
    opts = orc.dirk_opts(A=A, c=c, nstage=7, rtol=1e-13, atol=1e-16, newton_tol=1e-3,
                         max_iters=30, max_setups=5)

    def run(N):
        ch = orc.Chem(n)
        y = y0.copy()
        ch.prepare(y)
        ch.extra()
        for _ in range(N):
            failed, _ = orc.dirk_step(ch, opts, T / N, y)
            assert failed == 0
        return y.reshape(n, 10) * ch.array("scale").reshape(n, 10)

    ref = run(512)
    scale = np.abs(ref).max(axis=0)
    errs = [np.max(np.abs(run(N) - ref) / scale) for N in (16, 32, 64)]
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    assert all(13.0 < r < 19.0 for r in ratios), (errs, ratios)


def _prepared(n, seed, **kw):
    y, _ = problems.chem_cells(n, seed=seed, **kw)
    ch = orc.Chem(n)
    ys = y.copy()
    ch.prepare(ys)
    ch.extra()
    return y, ch, ys


def test_oracle_dirk_third_order():
    """Global error of N fixed SDIRK3 steps against N=512 steps: ratio ~ 8 per halving.
    Warm gas (3e3-8e3 K, n 0.1-10 cm^-3) over 1e8 s stays inside one table bin per cell,
    where the RHS is smooth."""
    n, T = 32, 1e8
    y0, _ = problems.chem_cells(n, seed=7, ntot=(1e-1, 1e1), temp=(3e3, 8e3))
    opts = orc.dirk_opts(rtol=1e-13, atol=1e-16, newton_tol=1e-3, max_iters=30, max_setups=5)

    def run(N):
        ch = orc.Chem(n)
        y = y0.copy()
        ch.prepare(y)
        ch.extra()
        for _ in range(N):
            failed, _ = orc.dirk_step(ch, opts, T / N, y)
            assert failed == 0
        return y.reshape(n, 10) * ch.array("scale").reshape(n, 10)

    ref = run(512)
    scale = np.abs(ref).max(axis=0)
    errs = [np.max(np.abs(run(N) - ref) / scale) for N in (16, 32, 64)]
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    assert all(6.0 < r < 9.5 for r in ratios), (errs, ratios)


def _invariants(yp):
    yp = yp.reshape(-1, 10)
    H = 2 * yp[:, 0] + 2 * yp[:, 1] + yp[:, 2] + yp[:, 3] + yp[:, 4]
    He = yp[:, 5] + yp[:, 6] + yp[:, 7]
    pos = yp[:, 3] + yp[:, 1] + yp[:, 6] + 2 * yp[:, 7]
    # the network changes de by exactly the change of net positive charge (f8 = f3 + f1 +
    # f6 + 2 f7 - f4 in its own units), whatever factor the initial condition used
    charge = yp[:, 8] - (pos - yp[:, 4])
    return H, He, charge, pos


@pytest.mark.parametrize("h", [1e3, 1e5])
def test_oracle_dirk_conserves_nuclei_and_charge(h):
    n = 2048
    y0, ch, ys = _prepared(n, seed=3)
    failed, iters = orc.dirk_step(ch, orc.dirk_opts(), h, ys)
    ok = iters >= 0
    assert ok.sum() == n - failed and ok.mean() > 0.9
    y1 = ys * ch.array("scale")
    H0, He0, c0, p0 = _invariants(y0)
    H1, He1, c1, _ = _invariants(y1)
    assert np.max(np.abs(H1 - H0)[ok] / H0[ok]) < 1e-13
    assert np.max(np.abs(He1 - He0)[ok] / He0[ok]) < 1e-13
    assert np.max(np.abs(c1 - c0)[ok] / p0[ok]) < 1e-12
    # the state did change (the invariants are not trivially kept)
    assert np.max(np.abs(y1 - y0).reshape(n, 10)[ok] / np.abs(y0).reshape(n, 10)[ok]) > 1e-6


def test_oracle_dirk_newton_failure_reported():
    n = 256
    _, ch, ys = _prepared(n, seed=5)
    failed, iters = orc.dirk_step(ch, orc.dirk_opts(max_iters=1, max_setups=0,
                                                    newton_tol=1e-30), 1e3, ys)
    assert failed == n and np.all(iters == -1)

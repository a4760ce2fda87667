"""IMEX additive Runge-Kutta evolution on the CPU oracle (the imex_chem_hydro driver),
pinned by: the third-order and coupling conditions of the ARK3(2)4L[2]SA pair (Kennedy &
Carpenter 2003) with its ESDIRK structure, and the fourth-order ones of ARK4(3)7L[2]SA
(Kennedy & Carpenter 2019, the reference's IMEX inputs' pair) with fourth-order convergence
of both of its halves; third-order convergence of the explicit half
alone (chemistry switched off, rhs_scale = 0: the implicit stages reduce to z = R) on a
smooth flow, and of the implicit half alone (a uniform gas at rest: fe() and the advection
terms vanish, so the step is the ESDIRK on the chemistry) against the same driver at a
step 16 times smaller; and conservation of the total mass with periodic boundaries."""
import importlib.util
import os

import numpy as np

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

UNITS = (3.0e70, 3.0857e30, 1.0e11)
EU = UNITS[0] / UNITS[1] / UNITS[2] / UNITS[2]


def test_ark324_table():
    Ae, Ai, b, c = orc.ark324_table().tables()
    e = np.ones(4)
    assert np.allclose(Ae @ e, c, rtol=0, atol=1e-15)
    assert np.allclose(Ai @ e, c, rtol=0, atol=1e-15)
    assert np.all(np.triu(Ae) == 0)
    assert np.all(np.triu(Ai, 1) == 0)
    g = Ai[1, 1]
    assert Ai[0, 0] == 0 and np.all(np.diag(Ai)[1:] == g)       # ESDIRK, one diagonal
    assert abs(g - 0.435866521508459) < 1e-15
    assert np.array_equal(b, Ai[3])                              # stiffly accurate
    for A in (Ae, Ai):
        assert abs(b @ e - 1) < 1e-15
        assert abs(b @ c - 1 / 2) < 1e-15
        assert abs(b @ c ** 2 - 1 / 3) < 1e-15
        assert abs(b @ A @ c - 1 / 6) < 1e-15


def setup_state(dims, seed=None, at_rest=False):
    n = dims[0] * dims[1] * dims[2]
    fluid, chem = problems.primordial_blast_mri(*dims, bump=3.0, width=0.25)
    if at_rest:
        # one uniform gas (the centre cell's state everywhere)
        c0 = n // 2 + dims[0] // 2
        fluid = [np.full(n, a[c0]) for a in fluid]
        chem = np.tile(chem.reshape(n, 10)[c0], n)
    ch = orc.Chem(n)
    y = chem.copy()
    ch.prepare(y)
    ch.extra()
    return ch, [a.copy() for a in fluid] + [y]


def run(dims, h, nsteps, rhs_scale, at_rest=False, bc=(0,) * 6, tab=None):
    ch, w = setup_state(dims, at_rest=at_rest)
    g = orc.make_grid(*dims, 10, bc, gamma=5 / 3)
    newton = orc.dirk_opts(rhs_scale=rhs_scale, rtol=1e-10, atol=1e-14, newton_tol=1e-3,
                           max_iters=20)
    rc = orc.imex_evolve(g, ch, EU, tab if tab is not None else orc.ark324_table(), newton, h,
                         nsteps, w)
    assert rc == 0
    return w


def errors(dims, T, steps, rhs_scale, at_rest, tab=None):
    ref = run(dims, T / (16 * steps[-1]), 16 * steps[-1], rhs_scale, at_rest, tab=tab)
    errs = []
    for m in steps:
        w = run(dims, T / m, m, rhs_scale, at_rest, tab=tab)
        errs.append(max(np.max(np.abs(w[q] - ref[q])) / np.max(np.abs(ref[q]))
                        for q in range(6) if np.max(np.abs(ref[q])) > 0))
    return np.array(errs)


def test_explicit_half_is_third_order():
    """Chemistry off (rhs_scale = 0): the implicit stages solve z = R exactly, and the
    driver is the explicit ARK3 on fexpl; errors fall by ~8 per halving."""
    errs = errors((8, 6, 6), 0.6, [4, 8, 16], 0.0, False)
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(rates > 2.6) and np.all(rates < 3.6), (errs, rates)


def test_implicit_half_is_third_order():
    """A uniform gas at rest: fexpl vanishes, the step is the ESDIRK on the chemistry."""
    ch, w = setup_state((4, 4, 4), at_rest=True)
    # fexpl of a uniform state at rest is zero on every field
    g = orc.make_grid(4, 4, 4, 10, (0,) * 6, gamma=5 / 3)
    flag, wd = orc.fslow(g, ch, EU, w)
    assert flag == 0
    assert all(np.all(a == 0) for a in wd[:4]) and np.all(wd[5] == 0)
    errs = errors((4, 4, 4), 2e-5, [2, 4, 8], UNITS[2], True)
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(rates > 2.6) and np.all(rates < 3.6), (errs, rates)


def _pair_conditions(Ae, Ai, b, c):
    """Residuals of the fourth-order conditions of an additive pair with shared b and c:
    the quadrature ones and, for every choice of the two tables in each position, the
    tree conditions b.Xc, b.(c*Xc), b.Xc^2, b.XYc."""
    r = [b.sum() - 1, b @ c - 1 / 2, b @ c ** 2 - 1 / 3, b @ c ** 3 - 1 / 4]
    for X in (Ae, Ai):
        r += [b @ (X @ c) - 1 / 6, b @ (c * (X @ c)) - 1 / 8, b @ (X @ c ** 2) - 1 / 12]
        for Y in (Ae, Ai):
            r.append(b @ (X @ (Y @ c)) - 1 / 24)
    return np.array(r)


def test_ark437_pair_table(mvd):
    """ARK4(3)7L[2]SA (the reference's IMEX inputs' etable 13 / itable 112): the explicit
    part strictly lower triangular, the implicit part the ESDIRK table, both rows summing to
    the shared nodes, all fourth-order conditions of the pair (coupling ones included) to
    round-off, and the library's table the same doubles; a wrong digit in one explicit entry
    breaks the conditions."""
    t = orc.ark437_pair()
    Ae, Ai, b, c = t.tables()
    assert np.all(np.triu(Ae) == 0) and np.all(np.triu(Ai, 1) == 0)
    assert np.allclose(Ae.sum(axis=1), c, rtol=0, atol=5e-16)
    assert np.allclose(Ai.sum(axis=1), c, rtol=0, atol=5e-16)
    assert np.array_equal(b, Ai[6])
    assert np.max(np.abs(_pair_conditions(Ae, Ai, b, c))) < 1e-15
    bad = Ae.copy()
    bad[5, 2] *= 1 + 1e-9
    bad[5, 0] -= bad[5].sum() - Ae[5].sum()     # (keeping the node)
    assert np.max(np.abs(_pair_conditions(bad, Ai, b, c))) > 1e-12
    d = mvd.ark437_pair()
    for f in ("Ae", "Ai", "b"):
        assert np.array_equal(np.array(getattr(d, f)[:]), np.array(getattr(t, f)[:])), f
    assert d.stages == 7


def test_ark437_explicit_half_is_fourth_order():
    """As the ARK3 test, with the seven-stage pair: errors fall by ~16 per halving."""
    errs = errors((8, 6, 6), 0.6, [4, 8, 16], 0.0, False, tab=orc.ark437_pair())
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(rates > 3.6) and np.all(rates < 4.8), (errs, rates)


def test_ark437_implicit_half_is_fourth_order():
    """A uniform gas at rest: the step is the pair's ESDIRK part on the chemistry (its
    explicit first stage included)."""
    errs = errors((4, 4, 4), 2e-5, [2, 4, 8], UNITS[2], True, tab=orc.ark437_pair())
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(rates > 3.6) and np.all(rates < 4.6), (errs, rates)


def test_imex_conserves_mass_with_periodic_boundaries():
    w0 = run((8, 8, 8), 1e-4, 0, UNITS[2])
    w = run((8, 8, 8), 1e-4, 3, UNITS[2])
    m0 = w0[0].sum()
    assert abs(w[0].sum() - m0) <= 1e-13 * m0

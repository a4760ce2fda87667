"""Multirate (MRI) evolution on the CPU oracle (configuration 5 path), pinned by: the
Knoth-Wolke slow table's third-order conditions (exact rationals, Knoth & Wolke 1998);
the fast-step count rule; the forced DIRK step reducing to y + h*force when the chemistry
is switched off (rhs_scale = 0), which places the forcing term; and conservation of the
total mass with periodic boundaries (the slow RHS telescopes, the fast stage leaves the
density to the forcing alone).
"""
import importlib.util
import math
import os

import numpy as np

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

UNITS = (3.0e70, 3.0857e30, 1.0e11)
EU = UNITS[0] / UNITS[1] / UNITS[2] / UNITS[2]


def test_kw3_table():
    c, A = orc.mri_kw3_table()
    b, As, cs = A[3, :3], A[:3, :3], c[:3]
    assert np.allclose(As.sum(axis=1), cs, atol=1e-16, rtol=0)
    assert c[3] == 1.0
    assert abs(b.sum() - 1) < 1e-15
    assert abs(b @ cs - 1 / 2) < 1e-15
    assert abs(b @ cs ** 2 - 1 / 3) < 1e-15
    assert abs(b @ As @ cs - 1 / 6) < 1e-15
    assert np.all(np.diff(c) > 0)            # every MIS stage integrates forward in time


def test_fast_step_count():
    c, _ = orc.mri_kw3_table()
    for h, hmax in ((0.05, 1e-4), (0.05, 0.01), (1.0, 3.0), (0.3, 0.1)):
        for i in (1, 2, 3):
            want = max(1, math.ceil((c[i] - c[i - 1]) * h / hmax))
            assert orc.lib().orc_mri_fast_steps(i, h, hmax) == want


def test_forced_step_without_chemistry_is_pure_forcing():
    n = 64
    y, _ = problems.chem_cells(n, seed=3)
    ch = orc.Chem(n)
    ch.prepare(y)
    ch.extra()
    rng = np.random.default_rng(4)
    force = rng.standard_normal(10 * n) * 1e-3
    y0 = y.copy()
    h = 0.7
    assert orc.dirk_step_forced(ch, orc.dirk_opts(rhs_scale=0.0), h, y, force) == 0
    assert np.allclose(y, y0 + h * force, rtol=1e-13, atol=1e-15)


def test_mri_conserves_mass_with_periodic_boundaries():
    dims = (8, 8, 8)
    n = 512
    fluid, chem = problems.primordial_blast_mri(*dims)
    g = orc.make_grid(*dims, 10, (0,) * 6, gamma=5 / 3)
    ch = orc.Chem(n)
    y = chem.copy()
    ch.prepare(y)
    ch.extra()
    w = [a.copy() for a in fluid] + [y]
    m0 = w[0].sum()
    assert orc.mri_evolve(g, ch, EU, orc.dirk_opts(rhs_scale=UNITS[2]), 0.05, 2, 0.01, w) == 0
    assert abs(w[0].sum() - m0) <= 1e-13 * m0

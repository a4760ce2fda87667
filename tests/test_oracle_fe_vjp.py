"""Pins of the oracle's fe() vector-Jacobian product (oracle/orc_vjp.cpp).

* one face: (dF/ds)^T Fbar against central differences of the face flux, every
  direction, with and without tracers;
* the whole fe(): the dot-product test <v, (fe(w+hu) - fe(w-hu))/2h> = <wbar, u> for random
  u and v, with periodic, mixed and reflecting boundaries (the ghost rules' signs and the
  E/N/F shifted copies enter through the scatter back to source cells);
* linearity in the cotangent.
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


def random_stencil(rng, nchem, direction):
    s = np.zeros((6, 5 + nchem))
    rho = 1.0 + 0.3 * rng.uniform(-1, 1, 6)
    vel = 0.4 * rng.standard_normal((6, 3))
    p = 1.0 + 0.5 * rng.uniform(-1, 1, 6)
    s[:, 0] = rho
    s[:, 1:4] = rho[:, None] * vel
    s[:, 4] = p / 0.4 + 0.5 * rho * (vel ** 2).sum(axis=1)
    if nchem:
        s[:, 5:] = 0.5 + 0.4 * rng.uniform(size=(6, nchem))
    return s


@pytest.mark.parametrize("nchem", [0, 2])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_face_flux_vjp_matches_central_differences(nchem, direction):
    rng = np.random.default_rng(10 * nchem + direction)
    for trial in range(4):
        s = random_stencil(rng, nchem, direction)
        nv = 5 + nchem
        Fb = rng.standard_normal(nv)
        sb = orc.face_flux_vjp(s, direction, Fb)
        fd = np.zeros_like(s)
        for l in range(6):
            for c in range(nv):
                h = 1e-7 * max(np.abs(s[:, c]).max(), 1e-3)     # relative to the field
                sp, sm = s.copy(), s.copy()
                sp[l, c] += h
                sm[l, c] -= h
                fd[l, c] = Fb @ (orc.face_flux(sp, direction) - orc.face_flux(sm, direction)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.max(np.abs(sb - fd)) <= 1e-6 * scale, (trial, np.max(np.abs(sb - fd)) / scale)


# (Dirichlet negates the density in the ghost layer, so its fe() is not finite; left out)
@pytest.mark.parametrize("bc,nchem", [((0,) * 6, 0), ((1, 3, 3, 1, 3, 1), 0),
                                      ((0,) * 6, 2), ((3, 1, 0, 0, 1, 3), 1)])
def test_fe_vjp_dot_product(bc, nchem):
    n = (9, 8, 7)
    g = orc.make_grid(*n, nchem, bc)
    w = problems.smooth_state(*n, nchem, seed=5)
    rng = np.random.default_rng(1)
    ncell = n[0] * n[1] * n[2]
    u = [rng.standard_normal(ncell) * np.abs(a).max() for a in w[:5]]
    u.append(rng.standard_normal(max(nchem, 1) * ncell) * (0.5 if nchem else 0.0))
    v = [rng.standard_normal(ncell) for _ in range(5)] + [rng.standard_normal(max(nchem, 1) * ncell)]
    if not nchem:
        v[5][:] = 0.0
    wbar = orc.fe_vjp(g, w, v)
    h = 1e-6
    wp = [a + h * b for a, b in zip(w, u)]
    wm = [a - h * b for a, b in zip(w, u)]
    fp = orc.fe(g, wp)[1]
    fm = orc.fe(g, wm)[1]
    lhs = sum(float(np.dot(vv, (a - b) / (2 * h))) for vv, a, b in zip(v, fp, fm))
    rhs = sum(float(np.dot(a, b)) for a, b in zip(wbar, u))
    mag = sum(float(np.dot(np.abs(a), np.abs(b))) for a, b in zip(wbar, u))
    assert abs(lhs - rhs) <= 1e-6 * mag, (lhs, rhs, mag)


def test_fe_vjp_linear_in_cotangent():
    n = (8, 7, 6)
    g = orc.make_grid(*n, 0, (0,) * 6)
    w = problems.smooth_state(*n, 0, seed=8)
    rng = np.random.default_rng(2)
    ncell = n[0] * n[1] * n[2]
    v1 = [rng.standard_normal(ncell) for _ in range(5)] + [np.zeros(1)]
    v2 = [rng.standard_normal(ncell) for _ in range(5)] + [np.zeros(1)]
    a, b = 0.7, -1.3
    l = orc.fe_vjp(g, w, [a * x + b * y for x, y in zip(v1, v2)])
    r1, r2 = orc.fe_vjp(g, w, v1), orc.fe_vjp(g, w, v2)
    for q in range(5):
        ref = a * r1[q] + b * r2[q]
        assert np.max(np.abs(l[q] - ref)) <= 1e-12 * np.abs(ref).max()

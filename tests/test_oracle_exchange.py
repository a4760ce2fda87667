"""Ghost exchange of the oracle (orc_exchange, orc_gather) against the reference's own
exchange test and its boundary fills.

* The index-coded known-answer test of communication_test_main.cpp:133-605: every value is
  0.TTsXXXyyyZZZ (task, field, global x, y, z), all boundaries periodic, one task (its own
  neighbour); receive buffers are checked slot by slot against the coded value of the cell
  they must hold (the reference's formulas: W holds x = i + ieW - 2 with ieW = nx - 1, E
  x = i + isE with isE = 0, likewise S/N in y and B/F in z), to the reference's 1e-12, and
  the pack1D stencils of interior cells against x = i + l - 3 (the test's second part).
* The physical-boundary fills of EulerData::ExchangeStart (euler3D.hpp:793-1166), read
  per face and per boundary type from the reference's loops: low faces W/S/B hold the
  mirrored cell 2 - s, high faces E/N/F the shifted copy n - 3 + s; Neumann copies,
  reflecting negates the face-normal momentum (mx on W/E, my on S/N, mz on B/F),
  Dirichlet negates every value including rho, et and the species. The coded values
  identify the source cell and field, and the sign the negation.
"""
import itertools

import numpy as np
import pytest

import orc

PER, NEU, DIR, REF = 0, 1, 2, 3


def code(v, X, Y, Z):
    """The test's sum, in its order: task + species + x + y + z (task 0)."""
    return 0.0 + 0.001 * v + 1e-6 * X + 1e-9 * Y + 1e-12 * Z


def coded_state(nx, ny, nz, nchem):
    """Fields of the exchange test: value = 0.001 v + 1e-6 X + 1e-9 Y + 1e-12 Z (task 0)."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    w = [np.ascontiguousarray(code(v, x, y, z).reshape(-1)) for v in range(5)]
    if nchem:
        c = np.stack([code(5 + v, x, y, z) for v in range(nchem)], axis=-1)
        w.append(np.ascontiguousarray(c.reshape(-1)))
    else:
        w.append(np.zeros(1))
    return w


def buffer_view(buf, nv, na, nb):
    """BUFINDX(v,s,a,b,NVAR,3,na,nb) = v + NVAR*(s + 3*(a + na*b)) -> [b][a][s][v]."""
    return buf.reshape(nb, na, 3, nv)


@pytest.mark.parametrize("dims,nchem", [((7, 6, 5), 0), ((8, 5, 6), 4), ((5, 7, 9), 2)])
def test_exchange_index_coded_kat(dims, nchem):
    nx, ny, nz = dims
    nv = 5 + nchem
    g = orc.make_grid(nx, ny, nz, nchem, (PER,) * 6)
    w = coded_state(nx, ny, nz, nchem)
    W, E, S, N, B, F = orc.exchange(g, w)
    tol = 1e-12
    ieW, isE, jeS, jsN, keB, ksF = nx - 1, 0, ny - 1, 0, nz - 1, 0
    Wv, Ev = buffer_view(W, nv, ny, nz), buffer_view(E, nv, ny, nz)
    Sv, Nv = buffer_view(S, nv, nx, nz), buffer_view(N, nv, nx, nz)
    Bv, Fv = buffer_view(B, nv, nx, ny), buffer_view(F, nv, nx, ny)
    errs = 0
    for k, j, i, v in itertools.product(range(nz), range(ny), range(3), range(nv)):
        errs += abs(Wv[k, j, i, v] - code(v, i + ieW - 2, j, k)) > tol
        errs += abs(Ev[k, j, i, v] - code(v, i + isE, j, k)) > tol
    for k, j, i, v in itertools.product(range(nz), range(3), range(nx), range(nv)):
        errs += abs(Sv[k, i, j, v] - code(v, i, j + jeS - 2, k)) > tol
        errs += abs(Nv[k, i, j, v] - code(v, i, j + jsN, k)) > tol
    for k, j, i, v in itertools.product(range(3), range(ny), range(nx), range(nv)):
        errs += abs(Bv[j, i, k, v] - code(v, i, j, k + keB - 2)) > tol
        errs += abs(Fv[j, i, k, v] - code(v, i, j, k + ksF)) > tol
    assert errs == 0
    # pack1D of the interior cells (the test's loops over 3 <= i < n - 2)
    for d in range(3):
        for k, j, i in itertools.product(range(3, nz - 2), range(3, ny - 2), range(3, nx - 2)):
            s = orc.gather(g, w, d, i, j, k)
            for l, v in itertools.product(range(6), range(nv)):
                X, Y, Z = i + (l - 3) * (d == 0), j + (l - 3) * (d == 1), k + (l - 3) * (d == 2)
                assert abs(s[l, v] - code(v, X, Y, Z)) <= tol


def expected_physical(face, bc, s, a, b, v, n):
    """Value the reference's fill writes to slot s of face at transverse (a, b)."""
    d, high = face // 2, face % 2 == 1
    src = n[d] - 3 + s if high else 2 - s
    if d == 0:
        X, Y, Z = src, a, b
    elif d == 1:
        X, Y, Z = a, src, b
    else:
        X, Y, Z = a, b, src
    val = code(v, X, Y, Z)
    if bc == DIR or (bc == REF and v == 1 + d):
        val = -val
    return val


@pytest.mark.parametrize("bc_type", [NEU, REF, DIR])
def test_exchange_physical_boundaries(bc_type):
    nx, ny, nz, nchem = 6, 7, 5, 2
    nv = 5 + nchem
    n = (nx, ny, nz)
    w = coded_state(nx, ny, nz, nchem)
    g = orc.make_grid(nx, ny, nz, nchem, (bc_type,) * 6)
    bufs = orc.exchange(g, w)
    for face in range(6):
        d = face // 2
        na = ny if d == 0 else nx
        nb = ny if d == 2 else nz
        view = buffer_view(bufs[face], nv, na, nb)
        for b, a, s, v in itertools.product(range(nb), range(na), range(3), range(nv)):
            assert view[b, a, s, v] == expected_physical(face, bc_type, s, a, b, v, n), \
                (face, s, a, b, v)


def test_exchange_mixed_boundaries_match_gather():
    """Buffers and stencil packs read the same ghost rules: the slots of each face equal the
    ghost points of the boundary cells' stencils (mixed boundary types, one per face)."""
    nx, ny, nz, nchem = 6, 5, 7, 1
    nv = 5 + nchem
    bcs = (NEU, DIR, PER, PER, DIR, REF)
    g = orc.make_grid(nx, ny, nz, nchem, bcs)
    rng = np.random.default_rng(3)
    w = [rng.standard_normal(nx * ny * nz) for _ in range(5)] + \
        [rng.standard_normal(nx * ny * nz * nchem)]
    W, E, S, N, B, F = orc.exchange(g, w)
    Wv, Ev = buffer_view(W, nv, ny, nz), buffer_view(E, nv, ny, nz)
    Bv, Fv = buffer_view(B, nv, nx, ny), buffer_view(F, nv, nx, ny)
    for k, j in itertools.product(range(nz), range(ny)):
        lo = orc.gather(g, w, 0, 0, j, k)            # points -3..2 of the lowest x-face
        hi = orc.gather(g, w, 0, nx - 1, j, k)       # points n-4..n+1 of the highest
        assert np.array_equal(lo[:3], Wv[k, j])
        assert np.array_equal(hi[4:], Ev[k, j, :2])  # points n, n+1
    for j, i in itertools.product(range(ny), range(nx)):
        lo = orc.gather(g, w, 2, i, j, 0)
        hi = orc.gather(g, w, 2, i, j, nz - 1)
        assert np.array_equal(lo[:3], Bv[j, i])
        assert np.array_equal(hi[4:], Fv[j, i, :2])

"""Pins the CPU oracle's fe() restatement without the reference build (which needs SUNDIALS and
RAJA, absent here): exact closed forms, invariants and an order-of-accuracy check that a
dropped term, wrong sign or transposed index in face_flux / the ghost rules would fail.
"""
import importlib.util
import os

import numpy as np
import pytest

import orc
from conftest import PKG, ROOT

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)
spec = importlib.util.spec_from_file_location("mvd_bench", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(spec)
spec.loader.exec_module(bench)

BCS = [(0,) * 6, (1,) * 6, (3,) * 6, (2,) * 6, (0, 0, 1, 3, 3, 1)]


def test_face_flux_opcount():
    """The op counts bench.py uses for the FP64 roofline are the oracle's counted ones."""
    for nchem, expect in bench.FACE_FLOPS.items():
        c = orc.face_flux_opcount(nchem)
        assert c["add"] + c["mul"] + c["div"] + c["sqrt"] == expect
    c0 = orc.face_flux_opcount(0)
    assert c0["sqrt"] == 8          # 2 Roe + 6 sound speeds (utilities.cpp:298-299,378)


@pytest.mark.parametrize("bc", [(0,) * 6, (1,) * 6, (3,) * 6])
def test_constant_state_gives_zero(bc):
    """A uniform state has equal fluxes on every face: fe == 0 exactly (compile_test.cpp idea).
    Reflecting walls need zero normal momentum to keep the state uniform."""
    g = orc.make_grid(7, 5, 6, 2, bc)
    n = 7 * 5 * 6
    m = (0.0, 0.0, 0.0) if bc[0] == 3 else (0.3, -0.2, 0.1)
    w = [np.full(n, 1.2), np.full(n, m[0]), np.full(n, m[1]), np.full(n, m[2]),
         np.full(n, 5.0), np.tile([0.4, 0.7], n)]
    flag, wd = orc.fe(g, w)
    assert flag == 0
    for a in wd:
        assert np.count_nonzero(a) == 0


def test_constant_stencil_face_flux_is_physical_flux():
    """For a constant stencil WENO returns the split flux exactly, so the face flux must equal
    the physical flux F(w): this pins LV as the inverse of RV (utilities.cpp:318-364), the
    split, the projection and back-projection, and the weight normalisation."""
    rng = np.random.default_rng(3)
    gam = 1.4
    for _ in range(50):
        rho = rng.uniform(0.2, 3.0)
        vel = rng.normal(0, 1.0, 3)
        p = rng.uniform(0.1, 5.0)
        m = rho * vel
        et = p / (gam - 1) + 0.5 * rho * vel @ vel
        c = rng.uniform(0.1, 2.0, 2)
        for d in range(3):
            st = np.tile(np.r_[rho, m, et, c], (6, 1))
            f = orc.face_flux(st, d, gam)
            u = vel[d]
            exact = np.r_[rho * u, m * u, u * (et + p), c * u]
            exact[1 + d] += p
            assert np.allclose(f, exact, rtol=1e-13, atol=1e-13 * np.abs(exact).max())


def _advect_error(d, N):
    dims = [3, 3, 3]
    dims[d] = N
    g = orc.make_grid(*dims, 0)
    zz, yy, xx = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]),
                             indexing="ij")
    xc = ((xx, yy, zz)[d] + 0.5) / N
    rho = 1 + 0.5 * np.sin(2 * np.pi * xc)
    m = [np.zeros(rho.shape) for _ in range(3)]
    m[d] = rho * 1.0
    et = 1.0 / 0.4 + 0.5 * rho
    w = [rho.ravel(), m[0].ravel(), m[1].ravel(), m[2].ravel(), et.ravel(), np.zeros(1)]
    w = [np.ascontiguousarray(a) for a in w]
    flag, wd = orc.fe(g, w)
    exact = -0.5 * 2 * np.pi * np.cos(2 * np.pi * xc)
    return np.abs(wd[0] - exact.ravel()).max()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_weno5_fifth_order(d):
    """Uniform velocity and pressure reduce the Euler equations to advection of density, whose
    exact derivative is known; FD-WENO5 must converge at 5th order (a wrong stencil coefficient,
    smoothness indicator or weight drops the order)."""
    errs = np.array([_advect_error(d, N) for N in (16, 32, 64)])
    rates = np.log2(errs[:-1] / errs[1:])
    assert rates.min() > 4.6, rates


def test_periodic_fe_sums_to_zero():
    """With periodic boundaries each face flux enters two cells with opposite signs."""
    dims = (10, 9, 8)
    w = problems.smooth_state(*dims, nchem=2, seed=2, noise=0.05)
    g = orc.make_grid(*dims, 2, (0,) * 6)
    flag, wd = orc.fe(g, w)
    for a in wd[:5]:
        assert abs(a.sum()) <= 1e-12 * np.abs(a).sum()


def test_direction_symmetry():
    """Relabelling x -> y (grid axes and momentum components together) relabels fe the same way
    (the reference routes every direction through one face_flux by swapping components)."""
    nx, ny, nz = 9, 7, 6
    w = problems.smooth_state(nx, ny, nz, seed=4, noise=0.03)
    g = orc.make_grid(nx, ny, nz, 0, (0, 0, 1, 1, 3, 3))
    flag, wd = orc.fe(g, w)
    # transposed problem: new x = old y, new y = old x
    def T(a):
        return np.ascontiguousarray(a.reshape(nz, ny, nx).transpose(0, 2, 1).reshape(-1))
    wt = [T(w[0]), T(w[2]), T(w[1]), T(w[3]), T(w[4]), np.zeros(1)]
    gt = orc.make_grid(ny, nx, nz, 0, (1, 1, 0, 0, 3, 3))
    flag2, wdt = orc.fe(gt, wt)
    for a, b in ((wd[0], wdt[0]), (wd[1], wdt[2]), (wd[2], wdt[1]), (wd[3], wdt[3]),
                 (wd[4], wdt[4])):
        assert np.allclose(T(a), b, rtol=0, atol=1e-13 * np.abs(a).max())


def _buffers_1d(line, n, bc_lo, bc_hi, normal_comp):
    """Receive buffers as EulerData::ExchangeStart fills them for one rank
    (euler3D.hpp:607-786 periodic self-exchange, 793-1166 physical boundaries).
    line: (nv, n) values along the line. Returns W[nv,3], E[nv,3]."""
    nv = line.shape[0]
    sgn_lo = np.ones(nv)
    sgn_hi = np.ones(nv)
    if bc_lo == 2:
        sgn_lo[:] = -1
    if bc_lo == 3:
        sgn_lo[normal_comp] = -1
    if bc_hi == 2:
        sgn_hi[:] = -1
    if bc_hi == 3:
        sgn_hi[normal_comp] = -1
    W = np.zeros((nv, 3))
    E = np.zeros((nv, 3))
    for s in range(3):
        if bc_lo == 0:
            W[:, s] = line[:, n - 3 + s]          # received from the east send buffer
        else:
            W[:, s] = sgn_lo * line[:, 2 - s]
        if bc_hi == 0:
            E[:, s] = line[:, s]                  # received from the west send buffer
        else:
            E[:, s] = sgn_hi * line[:, n - 3 + s]
    return W, E


def _pack_bdry(line, W, E, i, n):
    """pack1D_*_bdry slot selection (euler3D.hpp:1256-1290)."""
    nv = line.shape[0]
    out = np.zeros((6, nv))
    for l in range(3):
        out[l] = W[:, i + l] if i < 3 - l else line[:, i - 3 + l]
        out[l + 3] = E[:, i - n + l] if i > n - l - 1 else line[:, i + l]
    return out


@pytest.mark.parametrize("bc", BCS)
def test_ghost_rules_match_receive_buffers(bc):
    """Index-coded fields (communication_test_main.cpp idea): the oracle's direct ghost rule
    must reproduce the reference's receive-buffer mechanism for every stencil of every face."""
    nx, ny, nz, nchem = 5, 4, 6, 2
    g = orc.make_grid(nx, ny, nz, nchem, bc)
    n = nx * ny * nz
    idx = np.arange(n)
    w = [1000.0 * (v + 1) + idx for v in range(5)]
    w = [np.ascontiguousarray(a, dtype=np.float64) for a in w]
    chem = np.ascontiguousarray(
        (np.arange(nchem)[None, :] + 1) * 1e5 + idx[:, None], dtype=np.float64).reshape(-1)
    w.append(chem)
    nv = 5 + nchem
    full = np.zeros((nv, nz, ny, nx))
    for v in range(5):
        full[v] = w[v].reshape(nz, ny, nx)
    for c in range(nchem):
        full[5 + c] = chem.reshape(n, nchem)[:, c].reshape(nz, ny, nx)
    for d, ndim in ((0, nx), (1, ny), (2, nz)):
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    pos = [i, j, k]
                    for face in range(ndim + 1):
                        if pos[d] != (face if face < ndim else ndim - 1):
                            continue
                        p2 = list(pos)
                        p2[d] = face
                        if d == 0:
                            line = full[:, k, j, :]
                        elif d == 1:
                            line = full[:, k, :, i]
                        else:
                            line = full[:, :, j, i]
                        W, E = _buffers_1d(line, ndim, bc[2 * d], bc[2 * d + 1], 1 + d)
                        expect = _pack_bdry(line, W, E, face, ndim)
                        got = orc.gather(g, w, d, *p2)
                        assert np.array_equal(got, expect), (d, p2, bc)


def test_stability_closed_form():
    """stability(): for a uniform state alpha = |mx/rho| + c; the reference uses |mx/rho| for
    every direction (utilities.cpp:509), so y/z momentum must not change alpha."""
    g = orc.make_grid(4, 4, 4, 0)
    n = 64
    rho, mx, my, mz, p = 1.3, 0.4, 2.0, -3.0, 0.9
    gam = 1.4
    et = p / (gam - 1) + 0.5 * (mx ** 2 + my ** 2 + mz ** 2) / rho
    w = [np.full(n, v) for v in (rho, mx, my, mz, et)] + [np.zeros(1)]
    a = orc.max_wave_speed(g, w)
    assert a == pytest.approx(abs(mx / rho) + np.sqrt(gam * p / rho), rel=1e-14)
    assert orc.stability_dt(g, w, 0.3) == pytest.approx(0.3 * 0.25 / a, rel=1e-15)


def test_illegal_state_flag():
    g = orc.make_grid(4, 4, 4, 0)
    w = problems.smooth_state(4, 4, 4, seed=9)
    w[0][5] = 0.0
    flag, _ = orc.fe(g, w)
    assert flag == -1

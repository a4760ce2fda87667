"""The ARKODE callbacks (libmvd_arkode.so, include/mvd_arkode.h) driven the way ARKStep and
MRIStep call them, with SUNDIALS' vectors, matrices, linear solver and inner stepper played
by the tests' stand-in (tests/sundials_stub), against the CPU oracle and the C ABI they sit
on:

* fEuler / stability / fslow equal mvd_fe / mvd_stability / mvd_fslow bit for bit (those are
  oracle-tested in test_gpu_fe.py and test_gpu_chem.py);
* ffast (multirate_chem_hydro_main.cpp:898-952) = TimeUnits x RHS on the chemistry plus
  sum_i tau^i F_i on all six subvectors, fimpl = ffast without forcing, frhs = RHS: against
  the oracle RHS (1e-12 per species) and exact numpy forcing sums;
* Jfast / Jimpl = TimeUnits x the reference Jacobian, Jrhs = the Jacobian (oracle, 1e-12);
* SUNLinSol_RankLocalLS: DIRECT type, setup on A = I - gamma J as ARKLS forms it, solve on
  subvector 5 only (equal to mvd_chem_lsetup / lsolve bit for bit, residual 1e-13), the
  singular-block and missing-subvector flags;
* the inner stepper: Evolve with a constant MRI forcing against the oracle's forced DIRK
  steps and fluid updates (1e-10), FullRhs = ffast without forcing, Reset;
* PostprocessStep against its numpy restatement (bit for bit).
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import arkode_harness as H
import orc
from conftest import PKG, normwise_err

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

pytestmark = pytest.mark.gpu
UNITS = (3.0e70, 3.0857e30, 1.0e11)
EU = UNITS[0] / UNITS[1] / UNITS[2] / UNITS[2]
TU = UNITS[2]
DIMS = (12, 10, 8)
BC = (3,) * 6


class Setup:
    """A context on the multirate blast (chemistry prepared), its device state, the oracle
    chemistry on the same cells, and the adapters' user data."""

    def __init__(self, mvd, torch, dims=DIMS, bc=BC, bump=100.0):
        self.mvd, self.torch = mvd, torch
        self.n = dims[0] * dims[1] * dims[2]
        self.fluid, self.chem = problems.primordial_blast_mri(*dims, bump=bump)
        self.ctx = mvd.Context(mvd.make_params(*dims, 10, bc, gamma=5 / 3, cfl=0.3,
                                               units=UNITS))
        self.ctx.chem_setup(orc.load_tables(), self.n)
        self.w = [torch.from_numpy(a.copy()).to("cuda") for a in self.fluid] + \
                 [torch.from_numpy(self.chem.copy()).to("cuda")]
        self.ctx.chem_prepare(self.w[5])
        self.oc = orc.Chem(self.n)
        self.yo = self.chem.copy()
        self.oc.prepare(self.yo)
        self.oc.extra()
        self.data = mvd.ArkodeData(ctx=self.ctx.h.value, time_units=TU)
        self.data.fast = mvd.sdirk3(rhs_scale=TU)
        H.stub()                        # SUNDIALS' symbols first, then the adapters
        self.A = mvd.arkode_lib()

    def zeros(self):
        return [self.torch.full_like(a, 7.0) for a in self.w]    # junk the callbacks overwrite

    def ud(self):
        return C.c_void_p(C.addressof(self.data))

    def close(self):
        self.ctx.close()


def test_fEuler_stability_fslow_equal_the_abi(mvd, torch_gpu):
    torch = torch_gpu
    s = Setup(mvd, torch)
    print(f"[arkode] fEuler/stability/fslow n={s.n}", flush=True)
    w = H.ManyVector(s.w)
    out = s.zeros()
    wdot = H.ManyVector(out)
    assert s.A.mvd_arkode_fEuler(0.0, w.h, wdot.h, s.ud()) == 0
    ref = s.zeros()
    s.ctx.fe(s.w, ref)
    for q in range(6):
        assert torch.equal(out[q], ref[q]), q
    dt = C.c_double()
    assert s.A.mvd_arkode_stability(w.h, 0.0, C.byref(dt), s.ud()) == 0
    assert dt.value == s.ctx.stability(s.w)
    # fslow overwrites w.et and round-trips the chemistry: run both on copies
    wa = [a.clone() for a in s.w]
    wb = [a.clone() for a in s.w]
    va = H.ManyVector(wa)
    oa = s.zeros()
    da = H.ManyVector(oa)
    assert s.A.mvd_arkode_fslow(0.0, va.h, da.h, s.ud()) == 0
    ob = s.zeros()
    s.ctx.fslow(wb, ob)
    for q in range(6):
        assert torch.equal(oa[q], ob[q]) and torch.equal(wa[q], wb[q]), q
    # an illegal state is the reference's -1
    wa[0][5] = -1.0
    assert s.A.mvd_arkode_fEuler(0.0, va.h, da.h, s.ud()) == -1
    for v in (w, wdot, va, da):
        v.free()
    s.close()


def test_ffast_fimpl_frhs_match_oracle(mvd, torch_gpu):
    """Row a17: the chemistry RHS times TimeUnits on subvector 5, zero fluid, and MRIStep's
    polynomial forcing sum_i tau^i F_i added to all six subvectors (nforcing = 3)."""
    torch = torch_gpu
    s = Setup(mvd, torch)
    print(f"[arkode] ffast/fimpl/frhs n={s.n}", flush=True)
    ref_rhs = s.oc.rhs(s.yo.copy()).reshape(s.n, 10)       # from the 1000 K warm start
    Ts0 = s.ctx.chem_state("Ts", torch).clone()
    dT0 = s.ctx.chem_state("dTs_ge", torch).clone()
    rng = np.random.default_rng(5)
    F = [[rng.standard_normal(a.numel()) for a in s.w] for _ in range(3)]
    # chemistry forcing well below each species' RHS, so that taking it off again in the
    # check below costs no digits of the RHS
    smax = np.abs(TU * ref_rhs).max(axis=0)
    for f in F:
        f[5] = (f[5].reshape(s.n, 10) * (1e-3 * smax)).reshape(-1)
    Fd = [H.ManyVector([torch.from_numpy(x).to("cuda") for x in f]) for f in F]
    st = C.c_void_p()
    sunctx = H.stub().stub_context()
    assert s.A.mvd_inner_stepper_create(sunctx, C.byref(s.data), C.byref(st)) == 0
    assert s.data.stepper == st.value
    tshift, tscale, t = 0.4, 0.25, 0.55
    arr = (C.c_void_p * 3)(*[f.h for f in Fd])
    H.stub().stub_set_forcing(st, tshift, tscale, 3, arr)
    w = H.ManyVector(s.w)
    out = s.zeros()
    wdot = H.ManyVector(out)
    assert s.A.mvd_arkode_ffast(t, w.h, wdot.h, s.ud()) == 0
    tau = (t - tshift) / tscale
    c = [1.0, tau, tau * tau]
    for q in range(6):
        got = out[q].cpu().numpy()
        base = np.zeros_like(got) if q < 5 else None
        if q < 5:
            want = ((base + c[0] * F[0][q]) + c[1] * F[1][q]) + c[2] * F[2][q]
            assert np.array_equal(got, want), q
        else:
            forc = ((c[0] * F[0][5]) + c[1] * F[1][5]) + c[2] * F[2][5]
            chem = (got - forc).reshape(s.n, 10)
            for sp in range(10):
                assert normwise_err(chem[:, sp], TU * ref_rhs[:, sp]) <= 1e-12, sp
    # fimpl: no forcing (MPIManyVector); frhs: the chemistry vector alone, no TimeUnits
    for fn, scale in (("mvd_arkode_fimpl", TU), ("mvd_arkode_frhs", 1.0)):
        s.ctx.chem_state_set("Ts", Ts0)
        s.ctx.chem_state_set("dTs_ge", dT0)
        o = s.zeros()
        if fn == "mvd_arkode_frhs":
            yl, fl = H.Leaf(s.w[5]), H.Leaf(o[5])
            assert getattr(s.A, fn)(t, yl.h, fl.h, s.ud()) == 0
            yl.free()
            fl.free()
        else:
            od = H.ManyVector(o)
            assert getattr(s.A, fn)(t, w.h, od.h, s.ud()) == 0
            for q in range(5):
                assert int(torch.count_nonzero(o[q])) == 0
            od.free()
        got = o[5].cpu().numpy().reshape(s.n, 10)
        for sp in range(10):
            assert normwise_err(got[:, sp], scale * ref_rhs[:, sp]) <= 1e-12, (fn, sp)
    # FullRhs through the stepper: ffast with the forcing disabled
    s.ctx.chem_state_set("Ts", Ts0)
    s.ctx.chem_state_set("dTs_ge", dT0)
    o2 = s.zeros()
    d2 = H.ManyVector(o2)
    assert H.stub().stub_stepper_fullrhs(st, t, w.h, d2.h, 0) == 0
    assert s.data.disable_forcing == 0
    for q in range(5):
        assert int(torch.count_nonzero(o2[q])) == 0
    got = o2[5].cpu().numpy().reshape(s.n, 10)
    for sp in range(10):
        assert normwise_err(got[:, sp], TU * ref_rhs[:, sp]) <= 1e-12, sp
    # a vector with five subvectors is refused
    w5 = H.ManyVector(s.w[:5])
    assert s.A.mvd_arkode_ffast(t, w5.h, wdot.h, s.ud()) == -1
    for v in [w, wdot, d2, w5] + Fd:
        v.free()
    H.stub().MRIStepInnerStepper_Free(C.byref(st))
    H.stub().stub_context_free(sunctx)
    s.close()


def test_jfast_jimpl_jrhs_match_oracle(mvd, torch_gpu):
    """Row a18: calculate_jacobian_cvklu into the dense block matrix, times TimeUnits for
    Jfast / Jimpl (J *= TUnit, multirate_chem_hydro_main.cpp:980-982)."""
    torch = torch_gpu
    s = Setup(mvd, torch)
    print(f"[arkode] Jfast/Jimpl/Jrhs n={s.n}", flush=True)
    yo = s.yo.copy()
    s.oc.rhs(yo)                       # the rates the Jacobian reuses
    Jref = s.oc.jac(yo)                # [cell, row, col]
    f = torch.empty_like(s.w[5])
    s.ctx.chem_rhs(s.w[5], f)
    w = H.ManyVector(s.w)
    for fn, scale in (("mvd_arkode_Jfast", TU), ("mvd_arkode_Jimpl", TU),
                      ("mvd_arkode_Jrhs", 1.0)):
        J = torch.full((100 * s.n,), 3.0, dtype=torch.float64, device="cuda")
        M = H.BlockMatrix(J, s.n)
        if fn == "mvd_arkode_Jrhs":
            yl = H.Leaf(s.w[5])
            assert getattr(s.A, fn)(0.0, yl.h, None, M.h, s.ud(), None, None, None) == 0
            yl.free()
        else:
            assert getattr(s.A, fn)(0.0, w.h, None, M.h, s.ud(), None, None, None) == 0
        got = J.cpu().numpy().reshape(s.n, 10, 10).transpose(0, 2, 1)   # column-major blocks
        for r in range(10):
            for c in range(10):
                assert normwise_err(got[:, r, c], scale * Jref[:, r, c]) <= 1e-12, (fn, r, c)
        M.free()
    # a matrix whose block count differs from the grid is refused
    J = torch.zeros(100 * (s.n - 1), dtype=torch.float64, device="cuda")
    M = H.BlockMatrix(J, s.n - 1)
    assert s.A.mvd_arkode_Jfast(0.0, w.h, None, M.h, s.ud(), None, None, None) == -1
    M.free()
    w.free()
    s.close()


def test_rank_local_ls(mvd, torch_gpu):
    """Row a19: SUNLinSol_RankLocalLS ops on the ManyVector, subvector 5 only."""
    torch = torch_gpu
    s = Setup(mvd, torch)
    print(f"[arkode] RankLocalLS n={s.n}", flush=True)
    L = H.stub()
    sunctx = L.stub_context()
    x = [torch.full_like(a, 9.0) for a in s.w]
    b = [torch.randn(a.numel(), dtype=torch.float64, device="cuda") for a in s.w]
    xv, bv = H.ManyVector(x), H.ManyVector(b)
    assert s.A.mvd_SUNLinSol_RankLocalLS(xv.h, 6, C.byref(s.data), sunctx) is None
    S = s.A.mvd_SUNLinSol_RankLocalLS(xv.h, 5, C.byref(s.data), sunctx)
    assert S
    assert L.stub_ls_gettype(S) == 0                 # SUNLINEARSOLVER_DIRECT
    assert L.stub_ls_initialize(S) == 0
    f = torch.empty_like(s.w[5])
    s.ctx.chem_rhs(s.w[5], f)
    J = torch.empty(100 * s.n, dtype=torch.float64, device="cuda")
    s.ctx.chem_jac(s.w[5], J)
    gamma = 0.37 * TU
    eye = torch.eye(10, dtype=torch.float64, device="cuda").reshape(1, 100).repeat(s.n, 1)
    A = (eye - gamma * J.reshape(s.n, 100)).reshape(-1).contiguous()   # ARKLS: I - gamma J
    M = H.BlockMatrix(A, s.n)
    assert L.stub_ls_setup(S, M.h) == 0
    assert L.stub_ls_solve(S, M.h, xv.h, bv.h, 1e-6) == 0
    assert L.stub_ls_lastflag(S) == 0
    for q in range(5):                               # the fluid part is not touched
        assert bool((x[q] == 9.0).all()), q
    # the same as the C ABI's own setup (I - gamma J formed inside) and solve
    xr = torch.empty_like(b[5])
    assert s.ctx.chem_lsetup(gamma, J) == 0
    s.ctx.chem_lsolve(b[5], xr)
    assert torch.equal(x[5], xr)
    Ab = A.reshape(s.n, 10, 10).transpose(1, 2).contiguous()      # row-major blocks
    res = (Ab @ x[5].reshape(s.n, 10, 1)).reshape(s.n, 10) - b[5].reshape(s.n, 10)
    scale = (Ab.abs() @ x[5].abs().reshape(s.n, 10, 1)).reshape(s.n, 10) + b[5].reshape(s.n, 10).abs()
    assert float((res.abs() / scale).max()) <= 1e-13
    # a singular block: LU factorisation failure (recoverable), then solve refuses
    A2 = A.clone()
    A2.reshape(s.n, 100)[7] = 0.0
    M2 = H.BlockMatrix(A2, s.n)
    assert L.stub_ls_setup(S, M2.h) == 808            # SUNLS_LUFACT_FAIL
    assert L.stub_ls_lastflag(S) == 808
    assert L.stub_ls_solve(S, M2.h, xv.h, bv.h, 1e-6) == -809   # SUNLS_PACKAGE_FAIL_UNREC
    # vectors without the subvector: SUNLS_MEM_FAIL
    assert L.stub_ls_setup(S, M.h) == 0
    x5 = H.ManyVector(x[:5])
    assert L.stub_ls_solve(S, M.h, x5.h, bv.h, 1e-6) == -803
    assert L.stub_ls_free(S) == 0
    for v in (xv, bv, M, M2, x5):
        v.free()
    L.stub_context_free(sunctx)
    s.close()


def test_inner_stepper_evolve_matches_oracle(mvd, torch_gpu):
    """RankLocalStepper_Evolve over [t0, tout] with MRIStep's constant forcing: fast steps of
    at most hmax_fast, fluid y += dt F, chemistry by the forced DIRK step (TimeUnits RHS
    scale); against the oracle's forced step and numpy for the fluid."""
    torch = torch_gpu
    s = Setup(mvd, torch, bump=3.0)
    print(f"[arkode] inner stepper n={s.n}", flush=True)
    L = H.stub()
    sunctx = L.stub_context()
    st = C.c_void_p()
    assert s.A.mvd_inner_stepper_create(sunctx, C.byref(s.data), C.byref(st)) == 0
    s.data.hmax_fast = 0.004
    rng = np.random.default_rng(9)
    F = [rng.standard_normal(a.numel()) * 1e-3 * max(float(a.abs().max()), 1e-30) for a in s.w]
    F[5] = rng.standard_normal(s.w[5].numel()) * 1e-7
    Fd = H.ManyVector([torch.from_numpy(x).to("cuda") for x in F])
    arr = (C.c_void_p * 1)(Fd.h)
    L.stub_set_forcing(st, 0.0, 1.0, 1, arr)
    y = H.ManyVector(s.w)
    t0, tout = 0.1, 0.11                     # 3 fast steps of 0.01/3
    rc = L.stub_stepper_evolve(st, t0, tout, y.h)
    assert rc == 0
    assert s.data.fast_steps == 3 and s.data.tcur == tout
    dt = (tout - t0) / 3
    wo = [a.copy() for a in s.fluid] + [s.yo.copy()]
    for _ in range(3):
        for q in range(5):
            wo[q] = wo[q] + dt * F[q]
        assert orc.dirk_step_forced(s.oc, orc.dirk_opts(rhs_scale=TU), dt, wo[5], F[5]) == 0
    for q in range(5):
        assert np.array_equal(s.w[q].cpu().numpy(), wo[q]), q
    got = s.w[5].cpu().numpy().reshape(s.n, 10)
    ref = wo[5].reshape(s.n, 10)
    for sp in range(10):
        assert normwise_err(got[:, sp], ref[:, sp]) <= 1e-10, sp
    # Reset moves the stepper's time; a forcing that varies over the step is refused
    assert L.stub_stepper_reset(st, 0.5, y.h) == 0 and s.data.tcur == 0.5
    arr2 = (C.c_void_p * 2)(Fd.h, Fd.h)
    L.stub_set_forcing(st, 0.0, 1.0, 2, arr2)
    assert L.stub_stepper_evolve(st, 0.5, 0.51, y.h) == -1
    y.free()
    Fd.free()
    L.MRIStepInnerStepper_Free(C.byref(st))
    L.stub_context_free(sunctx)
    s.close()


def test_postprocess_step(mvd, torch_gpu):
    """PostprocessStep: et from the scaled gas energy and the momenta, chemistry x*s/s."""
    torch = torch_gpu
    s = Setup(mvd, torch)
    rng = np.random.default_rng(3)
    for q in (1, 2, 3):                            # some motion
        s.w[q].copy_(torch.from_numpy(rng.standard_normal(s.n) * 1e-3).to("cuda") * s.w[0])
    before = [a.cpu().numpy() for a in s.w]
    scale = s.ctx.chem_state("scale", torch).cpu().numpy()
    w = H.ManyVector(s.w)
    assert s.A.mvd_arkode_postprocess(0.0, w.h, s.ud()) == 0
    rho, mx, my, mz = before[:4]
    ch = before[5] * scale
    ge = ch.reshape(s.n, 10)[:, 9]
    et = ge * (1.0 / EU) + 0.5 / rho * (mx * mx + my * my + mz * mz)
    assert np.array_equal(s.w[4].cpu().numpy(), et)
    assert np.array_equal(s.w[5].cpu().numpy(), ch / scale)
    for q in range(4):
        assert np.array_equal(s.w[q].cpu().numpy(), before[q])
    w.free()
    s.close()

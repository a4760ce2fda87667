"""Batched and plain GPU SDIRK chemistry steps through the C ABI against the CPU oracle.

Both GPU paths and the oracle run the same per-cell algorithm (same RHS evaluation sequence,
same Newton updates and convergence test, same LU operation order, no FMA contraction), so
they agree far inside the 1e-8 relative bar for integrated states: the tolerance below is
1e-10, per species, relative to the largest magnitude of that species over the batch, on
the cells that converged. Failed cells (Newton budget exhausted; the integrator would reject
the step) must be the same number on every path.
"""
import importlib.util
import os

import numpy as np
import pytest

import orc
from conftest import PKG, normwise_err

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

pytestmark = pytest.mark.gpu
TOL = 1e-10


def setup(mvd, torch, n, seed):
    y, _ = problems.chem_cells(n, seed=seed)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    return oc, yo, ctx, yd


def compare(got, ref, ok):
    g = got.reshape(-1, 10)[ok]
    r = ref.reshape(-1, 10)[ok]
    return max(normwise_err(g[:, s], r[:, s]) for s in range(10))


@pytest.mark.parametrize("n,seed,h", [(4096, 1, 1e3), (1000, 2, 1e3), (1, 3, 1e3),
                                      (4096, 4, 1e5), (2048, 5, 1e8)])
def test_dirk_batched_and_plain_match_oracle(mvd, torch_gpu, n, seed, h):
    torch = torch_gpu
    print(f"[dirk parity] n={n} seed={seed} h={h}", flush=True)
    oc, yo, ctx, yd = setup(mvd, torch, n, seed)
    _, _, ctx_p, y_plain = setup(mvd, torch, n, seed)      # same start for the plain path
    failed_o, iters_o = orc.dirk_step(oc, orc.dirk_opts(), h, yo)
    ok = iters_o >= 0
    failed_b, st_b = ctx.dirk_step(yd, h)
    Ts_b = ctx.chem_state("Ts", torch).cpu().numpy()
    failed_p, st_p = ctx_p.dirk_step(y_plain, h, plain=True)
    assert failed_b == failed_o == failed_p
    assert st_b["failed_cells"] == failed_o
    yb = yd.cpu().numpy()
    yp = y_plain.cpu().numpy()
    assert compare(yb, yo, ok) <= TOL
    assert compare(yp, yo, ok) <= TOL
    assert np.max(np.abs(Ts_b[ok] - oc.array("Ts")[ok]) / oc.array("Ts")[ok]) <= TOL
    # the Newton work on the converged cells is the oracle's (iteration histogram)
    hist_o = np.bincount(np.minimum(iters_o[ok], 15), minlength=16)
    if failed_o == 0:
        assert st_b["newton_iters"] == int(iters_o.sum())
        assert st_p["newton_iters"] == int(iters_o.sum())
        assert list(hist_o) == st_b["hist"] == st_p["hist"]
    assert st_b["setups"] == st_p["setups"] >= n
    assert st_b["max_cell_iters"] == st_p["max_cell_iters"]


def test_dirk_four_stage_table(mvd, torch_gpu):
    """A 4-stage SDIRK table (constant diagonal 1/4, stiffly accurate): the batched kernel
    keeps three stage derivatives per cell in LDS, which for 256-thread blocks would exceed
    a CU's LDS, so the launch runs 128-thread blocks. Batched, plain and oracle agree as
    for the default table."""
    torch = torch_gpu
    n = 8192
    print(f"[dirk 4-stage] n={n}", flush=True)
    A = np.array([[0.25, 0, 0, 0], [0.25, 0.25, 0, 0], [0.2, 0.3, 0.25, 0],
                  [0.1, 0.3, 0.35, 0.25]])
    c = A.sum(axis=1)
    oc, yo, ctx, yd = setup(mvd, torch, n, 7)
    _, _, ctx_p, y_plain = setup(mvd, torch, n, 7)
    failed_o, iters_o = orc.dirk_step(oc, orc.dirk_opts(A=A, c=c, nstage=4), 1e3, yo)
    ok = iters_o >= 0
    o = mvd.set_table(mvd.sdirk3(), A, c)
    failed_b, st_b = ctx.dirk_step(yd, 1e3, o)
    failed_p, st_p = ctx_p.dirk_step(y_plain, 1e3, o, plain=True)
    assert failed_b == failed_o == failed_p
    assert compare(yd.cpu().numpy(), yo, ok) <= TOL
    assert np.array_equal(yd.cpu().numpy(), y_plain.cpu().numpy())
    assert st_b["hist"] == st_p["hist"]
    ctx.close()
    ctx_p.close()


@pytest.mark.parametrize("n,seed,h", [(4096, 51, 1e3), (2048, 52, 1e5), (1, 53, 1e3)])
def test_dirk_esdirk_batched_and_plain_match_oracle(mvd, torch_gpu, n, seed, h):
    """The reference inputs' implicit table, ARK4(3)7L[2]SA's ESDIRK part (seven stages, an
    explicit first stage): the batched kernel (its stage derivatives in the global buffer),
    the plain path and the oracle agree as for SDIRK3, with the same Newton work."""
    torch = torch_gpu
    print(f"[dirk esdirk] n={n} seed={seed} h={h}", flush=True)
    oc, yo, ctx, yd = setup(mvd, torch, n, seed)
    _, _, ctx_p, y_plain = setup(mvd, torch, n, seed)
    A, c = orc.ark437_table()
    failed_o, iters_o = orc.dirk_step(oc, orc.dirk_opts(A=A, c=c, nstage=7), h, yo)
    ok = iters_o >= 0
    o = mvd.ark437()
    failed_b, st_b = ctx.dirk_step(yd, h, o)
    failed_p, st_p = ctx_p.dirk_step(y_plain, h, o, plain=True)
    assert failed_b == failed_o == failed_p
    assert compare(yd.cpu().numpy(), yo, ok) <= TOL
    assert compare(y_plain.cpu().numpy(), yo, ok) <= TOL
    Ts_b = ctx.chem_state("Ts", torch).cpu().numpy()
    assert np.max(np.abs(Ts_b[ok] - oc.array("Ts")[ok]) / oc.array("Ts")[ok]) <= TOL
    if failed_o == 0:
        hist_o = np.bincount(np.minimum(iters_o, 15), minlength=16)
        assert st_b["newton_iters"] == st_p["newton_iters"] == int(iters_o.sum())
        assert list(hist_o) == st_b["hist"] == st_p["hist"]
    # the explicit stage takes no Newton iteration: six implicit stages per cell
    assert st_b["newton_iters"] >= 6 * (n - failed_o)
    ctx.close()
    ctx_p.close()


def test_dirk_esdirk_rejected_where_unsupported(mvd, torch_gpu):
    """The adjoint replays SDIRK tables of up to four stages only, and the allocation-free
    entry point has no room for the stage buffer of a seven-stage table."""
    torch = torch_gpu
    n = 64
    _, _, ctx, yd = setup(mvd, torch, n, 7)
    lam = torch.zeros_like(yd)
    lam0 = torch.empty_like(yd)
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_adjoint(yd, 100.0, 2, 1, lam, lam0, mvd.ark437())
    assert e.value.code == mvd.MVD_UNSUPPORTED
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step_noalloc(yd, 1e3, mvd.ark437())
    assert e.value.code == mvd.MVD_UNSUPPORTED
    ctx.close()


def test_dirk_reports_newton_failure(mvd, torch_gpu):
    torch = torch_gpu
    n = 777
    print(f"[dirk failure] n={n}", flush=True)
    _, _, ctx, yd = setup(mvd, torch, n, 6)
    o = mvd.sdirk3(max_iters=1, max_setups=0, newton_tol=1e-30)
    for plain in (False, True):
        y = yd.clone()
        failed, st = ctx.dirk_step(y, 1e3, opts=o, plain=plain)
        assert failed == n and st["failed_cells"] == n
        assert st["newton_iters"] == n and st["setups"] == n


def test_dirk_rejects_bad_options(mvd, torch_gpu):
    torch = torch_gpu
    _, _, ctx, yd = setup(mvd, torch, 64, 7)
    o = mvd.sdirk3()
    o.A[mvd.DIRK_MAXSTAGE + 1] = 0.3                  # unequal diagonal: not SDIRK
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step(yd, 1e3, opts=o)
    assert e.value.code == mvd.MVD_UNSUPPORTED
    o = mvd.ark437()
    o.A[3 * mvd.DIRK_MAXSTAGE + 3] = 0.2              # ESDIRK, unequal implicit diagonal
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step(yd, 1e3, opts=o)
    assert e.value.code == mvd.MVD_UNSUPPORTED
    o = mvd.sdirk3()
    o.A[1] = 0.1                                      # above the diagonal
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step(yd, 1e3, opts=o)
    assert e.value.code == mvd.MVD_BAD_ARGUMENT
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step(yd, -1.0)
    assert e.value.code == mvd.MVD_BAD_ARGUMENT


@pytest.mark.parametrize("table", ["sdirk3", "ark437"])
def test_dirk_tune_keeps_results(mvd, torch_gpu, table):
    """The autotuner times launch shapes on a scratch copy and restores the temperatures:
    a step after tuning equals a step without it (with the seven-stage table too, whose
    stage buffer is sized per shape)."""
    torch = torch_gpu
    n = 20000
    print(f"[dirk tune] n={n} table={table}", flush=True)
    _, _, ctx, yd = setup(mvd, torch, n, 8)
    _, _, ctx2, yd2 = setup(mvd, torch, n, 8)
    o = mvd.ark437() if table == "ark437" else mvd.sdirk3()
    ms = ctx.dirk_tune(yd, 1e3, o)
    assert ms > 0
    f1, s1 = ctx.dirk_step(yd, 1e3, o)
    f2, s2 = ctx2.dirk_step(yd2, 1e3, o)
    assert f1 == f2 == 0
    assert torch.equal(yd, yd2)
    assert s1["newton_iters"] == s2["newton_iters"]


@pytest.mark.parametrize("variant", ["plain", "forced", "substep", "esdirk"])
def test_dirk_work_queue_matches_default_shape(mvd, torch_gpu, variant):
    """8192 cells are 128 chunks of 64: on 16 resident waves each wave takes its first chunk
    by index and ~7 more from the atomic work queue (at the default one wave per SIMD every
    wave gets one chunk and the queue hands out none). The queue changes which wave solves
    a cell, not the cell's arithmetic: states, temperatures and Newton histograms equal the
    default shape's bit for bit and the oracle's to 1e-10, for the plain, forced and
    substepping kernels, with 64- and 256-thread blocks."""
    torch = torch_gpu
    n, seed, h = 8192, 31, 1e3
    print(f"[dirk queue] n={n} variant={variant}", flush=True)
    rng = np.random.default_rng(seed + 1)
    force = rng.standard_normal(10 * n) * 1e-5 if variant == "forced" else None
    mh = 4 if variant == "substep" else 0
    hstep = 1e5 if variant == "substep" else h
    kw = dict(max_iters=2, max_setups=0) if variant == "substep" else {}
    y, _ = problems.chem_cells(n, seed=seed)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    A7, c7 = orc.ark437_table()
    tab = dict(A=A7, c=c7, nstage=7) if variant == "esdirk" else {}
    oo = orc.dirk_opts(max_halvings=mh, **kw, **tab)
    if force is not None:
        fo = orc.dirk_step_forced(oc, oo, hstep, yo, force)
        iters_o = None
    else:
        fo, iters_o = orc.dirk_step(oc, oo, hstep, yo)
    runs = []
    for block, waves in ((256, 0), (256, 16), (64, 16), (128, 3)):
        ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
        ctx.chem_setup(orc.load_tables(), n)
        yd = torch.from_numpy(y.copy()).to("cuda")
        ctx.chem_prepare(yd)
        ctx.dirk_set_shape(block, waves)
        b, w = ctx.dirk_get_shape()
        assert b == block and (waves == 0 or w == waves)
        o = (mvd.ark437 if variant == "esdirk" else mvd.sdirk3)(max_halvings=mh, **kw)
        if force is not None:
            fd = ctx.dirk_step_forced(yd, hstep, torch.from_numpy(force).to("cuda"), o)
            st = None
        else:
            fd, st = ctx.dirk_step(yd, hstep, o)
        runs.append((fd, yd.cpu().numpy(), ctx.chem_state("Ts", torch).cpu().numpy(), st))
        ctx.close()
    f0, y0, T0, s0 = runs[0]
    assert f0 == fo
    for fd, yd_, Td, st in runs[1:]:
        assert fd == f0
        assert np.array_equal(yd_, y0) and np.array_equal(Td, T0)
        if st is not None:
            assert st["hist"] == s0["hist"] and st["newton_iters"] == s0["newton_iters"]
    ok = np.ones(n, dtype=bool) if iters_o is None else iters_o >= 0
    assert compare(y0, yo, ok) <= TOL
    if iters_o is not None and fo == 0:
        hist_o = np.bincount(np.minimum(iters_o, 15), minlength=16)
        assert list(hist_o) == s0["hist"]


def test_dirk_set_shape_rejects_bad_shapes(mvd, torch_gpu):
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    for block, waves in ((32, 0), (512, 0), (96, 0), (256, -1)):
        with pytest.raises(mvd.MvdError) as e:
            ctx.dirk_set_shape(block, waves)
        assert e.value.code == mvd.MVD_BAD_ARGUMENT
    ctx.close()


@pytest.mark.parametrize("variant", ["plain", "forced", "substep", "failing"])
def test_dirk_steps_equal_repeated_steps(mvd, torch_gpu, variant):
    """mvd_chem_dirk_steps runs nsteps steps per cell in one launch, each cell's steps back
    to back: states and temperatures equal nsteps single-step calls bit for bit, and the
    failed count is the sum of the single calls' (failed cell steps). `failing` takes a
    Newton budget so small that cells fail and carry their kept state into the next step;
    `substep` rejects and halves per cell; `forced` adds a constant forcing."""
    torch = torch_gpu
    n, seed, nsteps = 6000, 41, 5
    print(f"[dirk steps] n={n} variant={variant} nsteps={nsteps}", flush=True)
    rng = np.random.default_rng(seed)
    force = torch.from_numpy(rng.standard_normal(10 * n) * 1e-5).to("cuda") \
        if variant == "forced" else None
    kw = {"plain": {}, "forced": {}, "substep": dict(max_halvings=4, max_iters=2, max_setups=0),
          "failing": dict(max_iters=1, max_setups=0)}[variant]
    h = {"plain": 300.0, "forced": 300.0, "substep": 2e4, "failing": 1e5}[variant]
    y, _ = problems.chem_cells(n, seed=seed)
    res = []
    for fused in (False, True):
        ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
        ctx.chem_setup(orc.load_tables(), n)
        yd = torch.from_numpy(y.copy()).to("cuda")
        ctx.chem_prepare(yd)
        o = mvd.sdirk3(**kw)
        if fused:
            failed, st = ctx.dirk_steps(yd, h, nsteps, force, o)
        else:
            failed = 0
            for _ in range(nsteps):
                if force is not None:
                    failed += ctx.dirk_step_forced(yd, h, force, o)
                else:
                    failed += ctx.dirk_step(yd, h, o)[0]
        res.append((failed, yd.cpu().numpy(), ctx.chem_state("Ts", torch).cpu().numpy(),
                    ctx.chem_state("dTs_ge", torch).cpu().numpy()))
        ctx.close()
    (f0, y0, T0, d0), (f1, y1, T1, d1) = res
    if variant == "failing":
        assert f0 > 0
    else:
        assert np.all(np.isfinite(y0))
    assert f1 == f0
    assert np.array_equal(y1, y0) and np.array_equal(T1, T0) and np.array_equal(d1, d0)


@pytest.mark.parametrize("variant", ["plain", "forced", "substep", "failing", "esdirk-forced",
                                     "esdirk-substep"])
def test_dirk_steps_setup_reuse_matches_oracle(mvd, torch_gpu, variant):
    """ARKODE's lsetup frequency within one multi-step call (max_steps_between_setups = 7
    over 16 steps): cells set up at y_n only on their first step, after a failed or split
    step and every 7 steps since their last setup, and otherwise start Newton with the
    factors and rate estimate they hold. The batched launch against the oracle's multi-step
    call: states, temperatures, failed cell steps and (without failures) the per-cell Newton
    iterations in total. The same call with setups every step differs from it (the reuse
    is not a no-op)."""
    torch = torch_gpu
    n, seed, nsteps, msbp = 4096, 43, 16, 7
    print(f"[dirk steps reuse] n={n} variant={variant} nsteps={nsteps} msbp={msbp}", flush=True)
    esd = variant.startswith("esdirk-")
    variant = variant[len("esdirk-"):] if esd else variant
    A7, c7 = orc.ark437_table()
    tab = dict(A=A7, c=c7, nstage=7) if esd else {}
    lib_opts = mvd.ark437 if esd else mvd.sdirk3
    rng = np.random.default_rng(seed)
    force = rng.standard_normal(10 * n) * 1e-5 if variant == "forced" else None
    kw = {"plain": {}, "forced": {}, "substep": dict(max_halvings=4, max_iters=2, max_setups=0),
          "failing": dict(max_iters=1, max_setups=0)}[variant]
    h = {"plain": 300.0, "forced": 300.0, "substep": 2e4, "failing": 1e5}[variant]
    oc, yo, ctx, yd = setup(mvd, torch, n, seed)
    fd = torch.from_numpy(force).to("cuda") if force is not None else None
    failed_d, st = ctx.dirk_steps(yd, h, nsteps, fd,
                                  lib_opts(max_steps_between_setups=msbp, **kw))
    failed_o, iters_o = orc.dirk_steps(oc, orc.dirk_opts(max_steps_between_setups=msbp, **kw,
                                                         **tab), h, nsteps, yo, force)
    assert failed_d == failed_o
    if variant == "failing":
        assert failed_o > 0
    ok = np.ones(n, dtype=bool)
    assert compare(yd.cpu().numpy(), yo, ok) <= TOL
    Td = ctx.chem_state("Ts", torch).cpu().numpy()
    assert np.max(np.abs(Td - oc.array("Ts")) / oc.array("Ts")) <= TOL
    if failed_o == 0:
        assert st["newton_iters"] == int(iters_o.sum())
    # the oracle with a setup on every step takes another path (fewer reused factors)
    y1, _ = problems.chem_cells(n, seed=seed)
    oc1 = orc.Chem(n)
    oc1.prepare(y1)
    oc1.extra()
    orc.dirk_steps(oc1, orc.dirk_opts(max_steps_between_setups=1, **kw, **tab), h, nsteps, y1,
                   force)
    assert not np.array_equal(y1, yo)
    ctx.close()


def test_nvec_axpy_steps_equals_repeated_combinations(mvd, torch_gpu):
    torch = torch_gpu
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.standard_normal(100003)).to("cuda")
    y0 = torch.from_numpy(rng.standard_normal(100003) * 1e3).to("cuda")
    a, k = 1.7e-3, 37
    ya = y0.clone()
    ctx.nvec_axpy_steps(k, a, x, ya)
    yb = y0.clone()
    for _ in range(k):
        ctx.linear_combination([1.0, a], [yb, x], yb)
    assert torch.equal(ya, yb)
    ctx.close()

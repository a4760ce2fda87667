"""The benchmark configurations at their stated sizes, through the C ABI.

C2  256^3 blast, periodic, nchem = 0: the fast fe() kernels against the OpenMP oracle.
C3  64^3 = 262144 independent chemistry cells, one SDIRK3 step of 1e3 s with the chemistry
    input's Newton controls (tolerance 1e-3, 5 iterations): the batched kernel
    (default launch shape, which at this size takes chunks from the work queue) against the
    oracle, states to 1e-10 and the Newton histogram identical.
C4  128^3 with 10 species: dot-product tests of the fe() and chemistry vector-Jacobian
    products against central differences of the forward kernels on random directions.
C5  the multirate primordial blast with the configuration's options (reflecting walls, slow
    step min(CFL 0.3 bound, 0.05), fast steps of at most 1e-4 with per-cell step rejection,
    the input's Newton controls: tolerance 1e-2, 10 iterations, Jacobian setups at ARKODE's
    default frequency, every 20 fast steps):
    one slow step at 16^3 against the oracle, and two slow steps at 128^3 checked for a
    finite state, no failed cells and positive density and pressure.
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
UNITS = (3.0e70, 3.0857e30, 1.0e11)
EU = UNITS[0] / UNITS[1] / UNITS[2] / UNITS[2]
# Newton controls of the reference's inputs: input_primordial_ode.txt:73-74 (C3) and
# input_primordial_blast_mr.txt:70-71 (C5)
C3_NEWTON = {"newton_tol": 1e-3, "max_iters": 5}
C5_NEWTON = {"newton_tol": 1e-2, "max_iters": 10, "max_steps_between_setups": 20}


def test_c2_fe_256_matches_oracle(mvd, torch_gpu):
    torch = torch_gpu
    n = 256
    print(f"[C2] fe {n}^3 vs oracle", flush=True)
    w = problems.blast3d(n, seed=1234) + [np.zeros(1)]
    g = orc.make_grid(n, n, n, 0, (0,) * 6)
    orc.set_threads(orc.max_threads())
    flag, ref = orc.fe(g, w)
    assert flag == 0
    ctx = mvd.Context(mvd.make_params(n, n, n, 0))
    wd = [torch.from_numpy(a).to("cuda") for a in w[:5]] + [None]
    out = [torch.empty_like(a) for a in wd[:5]] + [None]
    ctx.fe(wd, out)
    for q in range(5):
        assert normwise_err(out[q].cpu().numpy(), ref[q]) <= 1e-12, q
    ctx.close()


def test_c3_batched_chemistry_262144_matches_oracle(mvd, torch_gpu):
    torch = torch_gpu
    n = 64 ** 3
    print(f"[C3] batched DIRK n={n} vs oracle", flush=True)
    y, _ = problems.chem_cells(n, seed=4321)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    orc.set_threads(orc.max_threads())
    failed_o, iters_o = orc.dirk_step(oc, orc.dirk_opts(**C3_NEWTON), 1e3, yo)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    block, waves = ctx.dirk_get_shape()
    assert (n + 63) // 64 > waves            # more chunks than waves: the queue runs
    failed_b, st = ctx.dirk_step(yd, 1e3, mvd.sdirk3(**C3_NEWTON))
    assert failed_o == failed_b == 0
    got = yd.cpu().numpy().reshape(n, 10)
    ref = yo.reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= 1e-10, s
    Tg = ctx.chem_state("Ts", torch).cpu().numpy()
    assert np.max(np.abs(Tg - oc.array("Ts")) / oc.array("Ts")) <= 1e-10
    hist_o = np.bincount(np.minimum(iters_o, 15), minlength=16)
    assert list(hist_o) == st["hist"]
    assert st["newton_iters"] == int(iters_o.sum())
    assert len(set(iters_o.tolist())) >= 4   # uneven Newton counts across cells
    ctx.close()


def c4_state(torch, n):
    """The configuration-4 fields: the 256^3 blast recipe at n^3 (seed 777) with 10
    prepared species (scaled, ~1) as tracers."""
    fluid = problems.blast3d(n, seed=777)
    y, _ = problems.chem_cells(n ** 3, seed=778)
    return fluid, y


def test_c4_fe_vjp_dot_product_128_nchem10(mvd, torch_gpu):
    """<v, (fe(w+hu) - fe(w-hu))/2h> = <(d fe/d w)^T v, u> over all 15 components at 128^3
    with the tracer kernels, cotangents and directions N(0,1) (directions scaled by each
    field's magnitude)."""
    torch = torch_gpu
    n = 128
    N = n ** 3
    print(f"[C4] fe vjp dot {n}^3 nchem=10", flush=True)
    fluid, y = c4_state(torch, n)
    ctx = mvd.Context(mvd.make_params(n, n, n, 10, (0,) * 6))
    ctx.chem_setup(orc.load_tables(), N)
    chem = torch.from_numpy(y).to("cuda")
    ctx.chem_prepare(chem)
    w = [torch.from_numpy(a).to("cuda") for a in fluid] + [chem]
    gen = torch.Generator(device="cuda").manual_seed(11)
    u = [torch.randn(a.numel(), dtype=torch.float64, device="cuda", generator=gen)
         * a.abs().max() for a in w]
    v = [torch.randn(a.numel(), dtype=torch.float64, device="cuda", generator=gen) for a in w]
    wb = [torch.empty_like(a) for a in w]
    ctx.fe_vjp(w, v, wb)
    h = 1e-6
    fp = [torch.empty_like(a) for a in w]
    fm = [torch.empty_like(a) for a in w]
    ctx.fe([a + h * b for a, b in zip(w, u)], fp)
    ctx.fe([a - h * b for a, b in zip(w, u)], fm)
    lhs = sum(float(torch.dot(vv, (a - b) / (2 * h))) for vv, a, b in zip(v, fp, fm))
    rhs = sum(float(torch.dot(a, b)) for a, b in zip(wb, u))
    mag = sum(float(torch.dot(a.abs(), b.abs())) for a, b in zip(wb, u))
    assert abs(lhs - rhs) <= 1e-6 * mag, (lhs, rhs, mag)
    # tracers alone: a direction only in the species must be seen by the tracer adjoint
    lhs_c = float(torch.dot(v[5], (fp[5] - fm[5]) / (2 * h)))
    assert abs(lhs_c) > 0
    ctx.close()


def test_c4_chem_vjp_dot_product_128(mvd, torch_gpu):
    """<v, (f(y+hu) - f(y-hu))/2h> = <(d f/d y)^T v, u> for the chemistry RHS at 128^3 cells
    (temperature re-solved from the same warm start on each side; the VJP is the exact
    derivative incl. the temperature root)."""
    torch = torch_gpu
    N = 128 ** 3
    print(f"[C4] chem vjp dot N={N}", flush=True)
    _, y = c4_state(torch, 128)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), N)
    yd = torch.from_numpy(y).to("cuda")
    ctx.chem_prepare(yd)
    gen = torch.Generator(device="cuda").manual_seed(12)
    yd *= 1.0 + 0.05 * torch.randn(yd.numel(), dtype=torch.float64, device="cuda",
                                   generator=gen)
    f = torch.empty_like(yd)
    for _ in range(3):                         # converge the warm starts
        ctx.chem_rhs(yd, f)
    Ts = ctx.chem_state("Ts", torch).clone()
    dT = ctx.chem_state("dTs_ge", torch).clone()
    u = torch.randn(yd.numel(), dtype=torch.float64, device="cuda", generator=gen) * yd.abs()
    v = torch.randn(yd.numel(), dtype=torch.float64, device="cuda", generator=gen)
    yb = torch.empty_like(yd)
    ctx.chem_vjp(yd, v, yb)
    h = 1e-7
    fp, fm = torch.empty_like(yd), torch.empty_like(yd)
    ctx.chem_state_set("Ts", Ts)
    ctx.chem_state_set("dTs_ge", dT)
    ctx.chem_rhs(yd + h * u, fp)
    ctx.chem_state_set("Ts", Ts)
    ctx.chem_state_set("dTs_ge", dT)
    ctx.chem_rhs(yd - h * u, fm)
    # per-cell terms, so that a few cells whose difference crosses a table node (the rates'
    # kinks) can be told from a wrong derivative: those are rare and bounded
    d_fd = (v * (fp - fm) / (2 * h)).reshape(N, 10).sum(1)
    d_vjp = (yb * u).reshape(N, 10).sum(1)
    mag = ((yb * u).abs()).reshape(N, 10).sum(1) + 1e-300
    rel = ((d_fd - d_vjp).abs() / mag).cpu().numpy()
    assert np.mean(rel > 1e-5) < 1e-3, np.sort(rel)[-20:]
    assert abs(float(d_fd.sum() - d_vjp.sum())) <= 1e-6 * float(mag.sum())
    ctx.close()


def c5_setup(mvd, torch, m):
    fluid, chem = problems.primordial_blast_mri(m)
    ctx = mvd.Context(mvd.make_params(m, m, m, 10, (3,) * 6, gamma=5 / 3, cfl=0.3,
                                      units=UNITS))
    ctx.chem_setup(orc.load_tables(), m ** 3)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
         [torch.from_numpy(chem.copy()).to("cuda")]
    ctx.chem_prepare(wd[5])
    return ctx, wd, fluid, chem


H0, HMAX = 0.05, 1e-4


@pytest.mark.parametrize("table", ["sdirk3", "ark437"])
def test_c5_multirate_16_one_slow_step_matches_oracle(mvd, torch_gpu, table):
    """One slow step at 16^3 with the input's options, fast stepper SDIRK3 or the input's
    itable 112 (ARK4(3)7L[2]SA's ESDIRK part): device and oracle agree to 1e-8."""
    torch = torch_gpu
    m = 16
    print(f"[C5] multirate {m}^3 one slow step vs oracle, {table}", flush=True)
    tab_o = {}
    if table == "ark437":
        A7, c7 = orc.ark437_table()
        tab_o = dict(A=A7, c=c7, nstage=7)
    make = mvd.ark437 if table == "ark437" else mvd.sdirk3
    ctx, wd, fluid, chem = c5_setup(mvd, torch, m)
    g = orc.make_grid(m, m, m, 10, (3,) * 6, gamma=5 / 3)
    dt_cfl = ctx.stability(wd)
    assert dt_cfl == orc.stability_dt(g, [a.copy() for a in fluid] + [chem.copy()], 0.3)
    h = min(dt_cfl, H0)
    fast = make(rhs_scale=UNITS[2], max_halvings=10, **C5_NEWTON)
    assert ctx.mri_evolve(wd, h, 1, HMAX, fast, consistent=True) == 0
    oc = orc.Chem(m ** 3)
    y = chem.copy()
    oc.prepare(y)
    oc.extra()
    wo = [a.copy() for a in fluid] + [y]
    orc.set_threads(orc.max_threads())
    assert orc.mri_evolve(g, oc, EU, orc.dirk_opts(rhs_scale=UNITS[2], max_halvings=10,
                                                   **C5_NEWTON, **tab_o), h, 1,
                          HMAX, wo, consistent=True) == 0
    for q in (0, 4):
        assert normwise_err(wd[q].cpu().numpy(), wo[q]) <= 1e-8, q
    for q in (1, 2, 3):                   # momenta: gas at rest, tiny and of either sign
        assert float(np.abs(wd[q].cpu().numpy() - wo[q]).max()) <= \
            1e-8 * max(float(np.abs(wo[q]).max()), 1e-300), q
    got = wd[5].cpu().numpy().reshape(-1, 10)
    ref = wo[5].reshape(-1, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= 1e-8, ("chem", s)
    ctx.close()


@pytest.mark.parametrize("table", ["sdirk3", "ark437"])
def test_c5_multirate_128_two_slow_steps_properties(mvd, torch_gpu, table):
    torch = torch_gpu
    m = 128
    print(f"[C5] multirate {m}^3 two slow steps, {table}", flush=True)
    ctx, wd, _, _ = c5_setup(mvd, torch, m)
    h = min(ctx.stability(wd), H0)
    make = mvd.ark437 if table == "ark437" else mvd.sdirk3
    fast = make(rhs_scale=UNITS[2], max_halvings=10, **C5_NEWTON)
    mass0 = float(wd[0].sum())
    assert ctx.mri_evolve(wd, h, 2, HMAX, fast, consistent=True) == 0
    for q in range(6):
        assert bool(torch.isfinite(wd[q]).all()), q
    rho, mx, my, mz, et = wd[:5]
    p = (5 / 3 - 1) * (et - 0.5 * (mx * mx + my * my + mz * mz) / rho)
    assert float(rho.min()) > 0 and float(p.min()) > 0
    # the walls see uniform gas at rest (the overdensity is ~exp(-200) there), so the
    # reflecting boundaries pass no mass and the total stays to rounding
    assert abs(float(wd[0].sum()) - mass0) <= 1e-10 * mass0
    ctx.close()

#!/usr/bin/env python3
"""Benchmark of the reacting-flow RHS hot path on one MI355X.

Headline (the one JSON line's "value"): fe() WENO5 flux divergence throughput in Gcell/s
on the 3D fluid blast at 256^3, FP64, periodic, nchem = 0, inputs resident in HBM.
One "step" is one fe() evaluation over the whole grid (mvd_fe: all kernels plus the
legal-state check and its host synchronisation).

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--n 256] [--no-cpu]

Also reported in the same line: the plain (reference-order) GPU kernels, per-kernel HIP
event timings with the roofline of the dominant kernel, and the CPU oracle timed on the
host on a bounded sample of the same workload.
"""
import argparse
import gc
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sundials-manyvector-demo_amd")

# Per-face operation count of the reference face_flux (utilities.cpp:270-479) for nchem=0:
# adds + muls + divides + square roots, counted by instantiating the oracle's restatement
# with an op-counting scalar (tests/test_oracle_fe.py::test_face_flux_opcount pins it).
FACE_FLOPS = {0: 1543, 10: 3273}
# divergence per cell and field: 3 differences, 3 divides, 2 adds, 1 subtract; the x and y
# parts (2 differences, 2 divides, 1 add) are formed by fe_xyface_div and the rest
# (1 difference, 1 divide, 1 add, 1 subtract) by the z kernel (fe_zface_div slot)
DIV_FLOPS_PER_VAR = 9
DIV_FLOPS_XY, DIV_FLOPS_Z = 5, 4
BYTES_PER_CELL_FE = 16 * 5          # read 5 fields + write 5 fields, FP64, nchem = 0
PEAK_FP64_TFLOPS = 78.6             # MI355X FP64 (vector and matrix) peak
# Newton controls of the reference's inputs (ARKODE's convergence test): the chemistry test
# (input_primordial_ode.txt:73-74: maxniters 5, nlconvcoef 0.001) for the batched
# chemistry step, the multirate blast (input_primordial_blast_mr.txt:70-71: maxniters 10,
# nlconvcoef 0.01) for the multirate run, whose fast steps set up their Jacobian at
# ARKODE's default frequency (msbp 20; the input leaves it)
CHEM_NEWTON = {"newton_tol": 1e-3, "max_iters": 5}
MRI_NEWTON = {"newton_tol": 1e-2, "max_iters": 10, "max_steps_between_setups": 20}
# input_primordial_blast_imex.txt:67-68 (maxniters, nlconvcoef)
IMEX_NEWTON = {"newton_tol": 1e-2, "max_iters": 10}
PEAK_HBM_GBS = 8000.0               # MI355X HBM3E


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def kernel_flops(name, n, nchem=0):
    """Algorithmic FP64 operations of one launch of each fe() kernel on an n^3 grid."""
    ff = FACE_FLOPS[nchem]
    nv = 5 + nchem
    faces = (n + 1) * n * n
    cells = n ** 3
    if name == "fe_xyface_div":
        return 2 * faces * ff + cells * nv * DIV_FLOPS_XY
    if name == "fe_zface_div":
        return faces * ff + cells * nv * DIV_FLOPS_Z
    if name == "fe_plain_faces":
        return 3 * faces * ff
    if name == "fe_plain_div":
        return cells * nv * DIV_FLOPS_PER_VAR
    raise KeyError(name)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "fe_256_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d["kernels"][kernel]["hbm_bytes_per_launch"]
    except (KeyError, ValueError):
        return None


def cpu_baseline(n_sample, threads):
    """The CPU oracle (a restatement of the reference fEuler) on the host: one fe() of the
    blast initial condition on an n_sample^3 grid, `threads` OpenMP threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    problems = _load("mvd_problems", os.path.join(PKG, "problems.py"))
    w = problems.blast3d(n_sample, seed=1234) + [np.zeros(1)]
    g = orc.make_grid(n_sample, n_sample, n_sample, 0, (0,) * 6)
    orc.set_threads(threads)
    t0 = time.perf_counter()
    flag, _ = orc.fe(g, w)
    dt = time.perf_counter() - t0
    assert flag == 0
    return n_sample ** 3 / dt / 1e9, dt


# FP64 operations of one cell's chemistry RHS (temperature solve, rate interpolation, the
# ten ODE right-hand sides) and Jacobian (calculate_rhs_cvklu / calculate_jacobian_cvklu,
# raja_primordial_network.cpp:1015-1965): additions, multiplications, divisions and square
# roots, counted by instantiating the reference's expressions with an op-counting scalar
# (tests/test_bench_contract.py pins them); transcendental calls are reported apart.
CHEM_RHS_FLOPS, CHEM_RHS_TRANSCENDENTALS, CHEM_JAC_FLOPS = 1207, 24, 638


def bench_chem(api, problems, torch, n, seed, steps, h=1e3):
    """Config-3 batched chemistry: one SDIRK3 step of dt = 1e3 s over n independent cells,
    batched persistent kernel vs the plain one-cell-per-thread kernels. Every timed step
    starts from the same state (y and the temperature warm start are restored outside the
    timed region); the device time of the step comes from HIP events around it."""
    y, _ = problems.chem_cells(n, seed=seed)
    ctx = api.Context(api.make_params(4, 4, 4, 10))
    ctx.chem_setup(api.load_tables(), n)
    y0 = torch.from_numpy(y).to("cuda")
    ctx.chem_prepare(y0)
    Ts0 = ctx.chem_state("Ts", torch)
    dT0 = ctx.chem_state("dTs_ge", torch)
    opts = api.sdirk3(**CHEM_NEWTON)
    yy = y0.clone()
    tuned_ms = ctx.dirk_tune(yy, h, opts)          # launch-shape autotuner (untimed)

    def run(plain, k):
        ms, st = [], None
        for _ in range(k):
            ctx.chem_state_set("Ts", Ts0)
            ctx.chem_state_set("dTs_ge", dT0)
            yy.copy_(y0)
            torch.cuda.synchronize()
            failed, st = ctx.dirk_step(yy, h, opts, plain=plain)
            assert failed == 0
            ms.append(st["ms"])
        return float(np.mean(ms)), st

    run(False, 2)
    run(True, 1)
    b_ms, b_st = run(False, steps)
    p_ms, p_st = run(True, max(2, steps // 2))
    assert b_st["newton_iters"] == p_st["newton_iters"] and b_st["hist"] == p_st["hist"]
    rhs_fl, jac_fl, rhs_trans = CHEM_RHS_FLOPS, CHEM_JAC_FLOPS, CHEM_RHS_TRANSCENDENTALS
    # RHS evaluations: one per Newton iteration plus one before each Jacobian setup
    nrhs = b_st["newton_iters"] + b_st["setups"]
    lu_fl = 2 * 10 ** 3 // 3 + 10 ** 2            # factor, per setup (approx.)
    solve_fl = 2 * 10 ** 2                        # triangular solves, per iteration
    flops = nrhs * rhs_fl + b_st["setups"] * (jac_fl + lu_fl) + b_st["newton_iters"] * solve_fl
    out = {
        "workload": f"{round(n ** (1 / 3))}^3 = {n} independent primordial-chemistry cells, "
                    f"one SDIRK3 step dt={h:g} s, Newton tol {CHEM_NEWTON['newton_tol']:g} / "
                    f"{CHEM_NEWTON['max_iters']} iterations, seed {seed}",
        "batched_ms": b_ms, "plain_ms": p_ms,
        "batched_cells_per_s": n / (b_ms * 1e-3), "plain_cells_per_s": n / (p_ms * 1e-3),
        "speedup_batched_vs_plain": p_ms / b_ms,
        "newton_iters": b_st["newton_iters"], "max_cell_iters": b_st["max_cell_iters"],
        "setups": b_st["setups"], "newton_hist": b_st["hist"],
        "rhs_evals": nrhs, "fp64_flops": flops,
        "batched_fp64_tflops": flops / (b_ms * 1e-3) / 1e12,
        "batched_fp64_frac": flops / (b_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
        "tuned_best_ms": tuned_ms,
        "op_count_per_cell": {"rhs_flops": rhs_fl, "rhs_transcendentals": rhs_trans,
                              "jac_flops": jac_fl},
    }
    ctx.close()
    return out


def bench_vjp(api, problems, torch, n, seed, reps):
    """Config-4 cost ratio t(VJP fe + VJP chem) / t(fe + chem RHS) at n^3 with 10 species:
    the blast fluid state (seed), the config-3 cell generator for the species (prepared,
    i.e. Dengo-scaled, so they are both the advected tracers of fe() and the chemistry
    state), cotangents ~ N(0,1). Each figure is the mean over `reps` synchronous calls,
    timed with HIP events on the null stream the library uses."""
    ncell = n ** 3
    fluid = problems.blast3d(n, seed=seed)[:5]
    y, _ = problems.chem_cells(ncell, seed=seed + 1)
    ctx = api.Context(api.make_params(n, n, n, 10, (0,) * 6))
    ctx.chem_setup(api.load_tables(), ncell)
    chem = torch.from_numpy(y).to("cuda")
    ctx.chem_prepare(chem)
    w = [torch.from_numpy(a.reshape(-1)).to("cuda") for a in fluid] + [chem]
    gen = torch.Generator(device="cuda").manual_seed(seed + 2)
    v = [torch.randn(a.numel(), dtype=torch.float64, device="cuda", generator=gen) for a in w]
    wd = [torch.empty_like(a) for a in w]
    wb = [torch.empty_like(a) for a in w]
    f = torch.empty_like(chem)
    yb = torch.empty_like(chem)
    for _ in range(3):                      # temperature warm start, allocations
        ctx.chem_rhs(chem, f)

    def tm(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def chem_rhs():
        ctx.chem_rhs(chem, f)

    def chem_vjp():
        ctx.chem_vjp(chem, v[5], yb)

    t_fe = tm(lambda: ctx.fe(w, wd))
    t_rhs = tm(chem_rhs)
    t_fevjp = tm(lambda: ctx.fe_vjp(w, v, wb))
    t_chvjp = tm(chem_vjp)
    # sensitivity of the domain-integrated gas energy (sum of mdensity * ge) to the initial
    # species through 50 fast chemistry substeps (SDIRK3, h = 100 s), checkpoint every 10
    md = ctx.chem_state("mdensity", torch)
    sc = ctx.chem_state("scale", torch).reshape(ncell, 10)
    lam = torch.zeros(ncell, 10, dtype=torch.float64, device="cuda")
    lam[:, 9] = md * sc[:, 9]
    lam = lam.reshape(-1)
    lam0 = torch.empty_like(lam)
    ysave = chem.clone()
    Ts0 = ctx.chem_state("Ts", torch).clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    failed = ctx.dirk_adjoint(chem, 100.0, 50, 10, lam, lam0)
    e1.record()
    torch.cuda.synchronize()
    t_sens = e0.elapsed_time(e1)
    chem.copy_(ysave)
    ctx.chem_state_set("Ts", Ts0)
    opts = api.sdirk3()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(50):
        ctx.dirk_step(chem, 100.0, opts)
    e1.record()
    torch.cuda.synchronize()
    t_fwd50 = e0.elapsed_time(e1)
    ctx.close()
    return {
        "workload": f"{n}^3 blast fluid (seed {seed}) with 10 prepared primordial species "
                    f"(seed {seed + 1}), periodic, cotangents N(0,1)",
        "fe_ms": t_fe, "chem_rhs_ms": t_rhs, "fe_vjp_ms": t_fevjp, "chem_vjp_ms": t_chvjp,
        "fe_vjp_over_fe": t_fevjp / t_fe, "chem_vjp_over_rhs": t_chvjp / t_rhs,
        "cost_ratio": (t_fevjp + t_chvjp) / (t_fe + t_rhs),
        "target": "<= 3",
        "sensitivity": {
            "what": "d(sum_cells mdensity*ge)/d(initial species) through 50 SDIRK3 chemistry "
                    "substeps of 100 s, checkpoint every 10, exact stage Jacobians",
            "ms": t_sens, "failed_cells": failed,
            "forward_50_batched_steps_ms": t_fwd50,
            "adjoint_over_forward": t_sens / t_fwd50,
            "max_abs_sensitivity": float(lam0.abs().max()),
        },
    }


def bench_fe_tracers(api, problems, torch, n, seed, reps):
    """fe() with 10 advected species (the RHS of configurations 4 and 5) at n^3: the
    configuration-4 fields (blast fluid, prepared species as tracers), periodic; the mean
    of `reps` calls and the per-kernel HIP-event times."""
    ncell = n ** 3
    fluid = problems.blast3d(n, seed=seed)[:5]
    y, _ = problems.chem_cells(ncell, seed=seed + 1)
    ctx = api.Context(api.make_params(n, n, n, 10, (0,) * 6))
    ctx.chem_setup(api.load_tables(), ncell)
    chem = torch.from_numpy(y).to("cuda")
    ctx.chem_prepare(chem)
    w = [torch.from_numpy(a.reshape(-1)).to("cuda") for a in fluid] + [chem]
    wd = [torch.empty_like(a) for a in w]
    for _ in range(3):
        ctx.fe(w, wd)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.fe(w, wd)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    prof = ctx.profile_read()
    ctx.close()
    kernels, total_fl, total_ms = {}, 0, 0.0
    for name, (kms, cnt) in prof.items():
        if cnt == 0:
            continue
        avg = kms / cnt
        fl = kernel_flops(name, n, 10)
        kernels[name] = {"avg_ms": avg, "launches": cnt, "tflops": fl / (avg * 1e-3) / 1e12,
                         "fp64_frac": fl / (avg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
        total_fl += fl
        total_ms += avg
    return {"workload": f"fe() {n}^3, 10 tracers (blast seed {seed}, species seed {seed + 1}), "
                        f"periodic", "ms_per_call": ms, "gcells": ncell / (ms * 1e-3) / 1e9,
            "kernels": kernels, "kernel_sum_ms": total_ms,
            "fp64_tflops": total_fl / (total_ms * 1e-3) / 1e12,
            "fp64_frac": total_fl / (total_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
            "hbm_algorithmic_gbs": 16 * 15 * ncell / (total_ms * 1e-3) / 1e9}


def bench_sod(api, problems, torch, cpu=True):
    """Config-1 Sod shock tube: 128x3x3 cells, Neumann, 2000 fixed steps of 1e-4 to t=0.2 with
    Bogacki-Shampine 3 (mvd_erk_evolve, one captured step replayed); the oracle's ERK on one
    host core for the same run, and the largest field-relative difference between the two."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    nx, nsteps, h = 128, 2000, 1e-4
    w0 = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in problems.sod(nx)[:5]]
    ctx = api.Context(api.make_params(nx, 3, 3, 0, (1,) * 6))
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in w0] + [None]
    ctx.erk_evolve(wd, h, 10)                     # capture the step graph
    for q in range(5):
        wd[q].copy_(torch.from_numpy(w0[q]))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    rc = ctx.erk_evolve(wd, h, nsteps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    out = {"workload": f"Sod {nx}x3x3, Neumann, ERK3 (Bogacki-Shampine) h={h:g}, {nsteps} steps "
                       f"to t=0.2",
           "rc": rc, "gpu_ms": ms, "steps_per_s": nsteps / (ms * 1e-3),
           "cell_steps_per_s": nx * 9 * nsteps / (ms * 1e-3)}
    if cpu:
        orc.set_threads(1)
        wc = w0 + [np.zeros(1)]
        wc = [a.copy() for a in wc]
        t0 = time.perf_counter()
        assert orc.erk_evolve(orc.make_grid(nx, 3, 3, 0, (1,) * 6),
                              orc.erk(*orc.bs3_table()), h, nsteps, wc) == 0
        out["cpu_1core_s"] = time.perf_counter() - t0
        diffs = []
        for q in range(5):
            den = float(np.abs(wc[q]).max())
            num = float(np.abs(wd[q].cpu().numpy() - wc[q]).max())
            diffs.append(num / den if den > 0 else num)      # my = mz = 0 stay zero
        out["max_rel_diff_vs_oracle"] = max(diffs)
    ctx.close()
    return out


def bench_mri(api, problems, torch, n, nslow, cpu=True, cpu_n=16, table="ark437"):
    """Config-5 multirate primordial blast: n^3 cells, 10 species, reflecting boundaries,
    gas at rest with a x100 Gaussian overdensity of width 0.05 at 1e4 K in 10 K gas (the
    reference's species recipe), units of input_primordial_blast_mr.txt. Slow step =
    min(CFL 0.3 bound, the input's h0 = 0.05); fast steps no longer than the input's
    hmax = 1e-4 (501 per slow step); fast stepper: the input's implicit table (itable 112,
    ARK4(3)7L[2]SA's ESDIRK part; table="sdirk3" for Alexander's SDIRK3) with ffast's
    TimeUnits scaling and per-cell step rejection; chemistry forcing converted to the scaled
    fast variables. The oracle runs the same options on a cpu_n^3 sample for one slow step
    (host baseline), and the device is compared with it there."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    units = (3.0e70, 3.0857e30, 1.0e11)
    eu = units[0] / units[1] / units[2] / units[2]
    h0, hmax = 0.05, 1e-4

    def setup(m):
        fluid, chem = problems.primordial_blast_mri(m)
        ctx = api.Context(api.make_params(m, m, m, 10, (3,) * 6, gamma=5 / 3, cfl=0.3,
                                          units=units))
        ctx.chem_setup(api.load_tables(), m ** 3)
        wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
             [torch.from_numpy(chem.copy()).to("cuda")]
        ctx.chem_prepare(wd[5])
        return ctx, wd, fluid, chem

    make = api.ark437 if table == "ark437" else api.sdirk3
    fast = make(rhs_scale=units[2], max_halvings=10, **MRI_NEWTON)
    tname = "ARK4(3)7L[2]SA-ESDIRK" if table == "ark437" else "SDIRK3"
    ctx, wd, _, _ = setup(n)
    h = min(ctx.stability(wd), h0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    failed = ctx.mri_evolve(wd, h, nslow, hmax, fast, consistent=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ok = all(bool(torch.isfinite(a).all()) for a in wd)
    ctx.close()
    nfast = sum(api.mri_fast_steps(i, h, hmax) for i in (2, 3, 4))
    out = {"workload": f"multirate primordial blast {n}^3, 10 species, reflecting, "
                       f"{nslow} slow steps of {h:g} (MIS Knoth-Wolke 3), {nfast} {tname} fast "
                       f"steps each (hmax {hmax:g}, Newton tol {MRI_NEWTON['newton_tol']:g} / "
                       f"{MRI_NEWTON['max_iters']} iterations, a Jacobian setup every "
                       f"{MRI_NEWTON['max_steps_between_setups']} fast steps), "
                       f"consistent forcing units",
           "slow_step": h, "fast_steps_per_slow_step": nfast, "failed_fast_cell_steps": failed,
           "finite": ok, "gpu_ms": ms,
           "cell_slow_steps_per_s": n ** 3 * nslow / (ms * 1e-3)}
    if cpu:
        m = cpu_n
        ctx, wd, fluid, chem = setup(m)
        hs = min(ctx.stability(wd), h0)
        ctx.mri_evolve(wd, hs, 1, hmax, fast, consistent=True)
        oc = orc.Chem(m ** 3)
        y = chem.copy()
        oc.prepare(y)
        oc.extra()
        wo = [a.copy() for a in fluid] + [y]
        g = orc.make_grid(m, m, m, 10, (3,) * 6, gamma=5 / 3)
        tab_o = {}
        if table == "ark437":
            A7, c7 = orc.ark437_table()
            tab_o = dict(A=A7, c=c7, nstage=7)
        orc.set_threads(1)
        t0 = time.perf_counter()
        orc.mri_evolve(g, oc, eu, orc.dirk_opts(rhs_scale=units[2], max_halvings=10, **MRI_NEWTON,
                                                **tab_o), hs, 1, hmax, wo, consistent=True)
        dt = time.perf_counter() - t0
        diffs = []
        for q in range(6):
            den = float(np.abs(wo[q]).max())
            num = float(np.abs(wd[q].cpu().numpy() - wo[q]).max())
            diffs.append(num / den if den > 0 else num)
        ctx.close()
        out["cpu_baseline"] = {"value": m ** 3 / dt, "unit": "cell*slow-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle multirate run, {m}^3 cells, 1 slow step, same "
                                         f"options, {dt:.1f} s"}
        out["sample_max_rel_diff_vs_oracle"] = max(diffs)
    return out


def cpu_chem_baseline(n_sample, seed, threads, reps, h=1e3):
    """The CPU oracle's per-cell SDIRK3 step on config-3 cells, `reps` times from the same
    initial state; returns cells/s and the total seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    problems = _load("mvd_problems", os.path.join(PKG, "problems.py"))
    y, _ = problems.chem_cells(n_sample, seed=seed)
    orc.set_threads(threads)
    total = 0.0
    for _ in range(reps):
        ch = orc.Chem(n_sample)
        ys = y.copy()
        ch.prepare(ys)
        ch.extra()
        t0 = time.perf_counter()
        failed, _ = orc.dirk_step(ch, orc.dirk_opts(**CHEM_NEWTON), h, ys)
        total += time.perf_counter() - t0
        assert failed == 0
        del ch
    return n_sample * reps / total, total


def quiet_gc():
    """Collect now and keep Python's cyclic collector out of the timed regions: with torch
    loaded a full collection walks several hundred thousand objects (~35 ms on the GPU
    box's host) and, left on, lands inside a timed loop about once per few hundred calls,
    which with 100 fe() calls of ~2.5 ms adds 0.35 ms to the mean. Reference counting still
    frees everything the bench drops."""
    gc.collect()
    gc.disable()


def bench_imex(api, problems, torch, n, nsteps, cpu=True, cpu_n=10, cpu_steps=500):
    """IMEX reacting flow (the imex_chem_hydro driver): the primordial blast of the multirate
    configuration on n^3 cells, reflecting boundaries, the units, pair and controls of
    input_primordial_blast_imex.txt (ARK4(3)7L[2]SA: etable 13 / itable 112; Newton 0.01 /
    10 iterations; rtol 1e-5, atol 1e-9), fixed steps of the input's hmax = 1e-4 where the
    reference adapts below it, chemistry forcing in the scaled variables. The oracle runs the
    same options on the input's own 10^3 grid for cpu_steps steps (host baseline), and the
    device is compared with it there."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    units = (3.0e70, 3.0857e30, 1.0e11)
    eu = units[0] / units[1] / units[2] / units[2]
    h = 1e-4

    def setup(m):
        fluid, chem = problems.primordial_blast_mri(m)
        ctx = api.Context(api.make_params(m, m, m, 10, (3,) * 6, gamma=5 / 3, units=units))
        ctx.chem_setup(api.load_tables(), m ** 3)
        wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
             [torch.from_numpy(chem.copy()).to("cuda")]
        ctx.chem_prepare(wd[5])
        return ctx, wd, fluid, chem

    newton = api.sdirk3(rhs_scale=units[2], **IMEX_NEWTON)
    tab = api.ark437_pair()
    ctx, wd, _, _ = setup(n)
    ctx.imex_evolve(wd, h, 1, newton, table=tab, consistent=True)     # allocations
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    failed = ctx.imex_evolve(wd, h, nsteps, newton, table=tab, consistent=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ok = all(bool(torch.isfinite(a).all()) for a in wd)
    ctx.close()
    out = {"workload": f"IMEX primordial blast {n}^3, 10 species, reflecting, {nsteps} steps of "
                       f"{h:g} with ARK4(3)7L[2]SA (etable 13 / itable 112), Newton tol "
                       f"{IMEX_NEWTON['newton_tol']:g} / {IMEX_NEWTON['max_iters']} iterations, "
                       f"consistent forcing units (after one untimed step)",
           "failed_cell_solves": failed, "finite": ok, "gpu_ms": ms,
           "cell_steps_per_s": n ** 3 * nsteps / (ms * 1e-3)}
    if cpu:
        m = cpu_n
        ctx, wd, fluid, chem = setup(m)
        ctx.imex_evolve(wd, h, cpu_steps, newton, table=tab, consistent=True)
        oc = orc.Chem(m ** 3)
        y = chem.copy()
        oc.prepare(y)
        oc.extra()
        wo = [a.copy() for a in fluid] + [y]
        g = orc.make_grid(m, m, m, 10, (3,) * 6, gamma=5 / 3)
        orc.set_threads(1)
        t0 = time.perf_counter()
        orc.imex_evolve(g, oc, eu, orc.ark437_pair(), orc.dirk_opts(rhs_scale=units[2],
                                                                    **IMEX_NEWTON),
                        h, cpu_steps, wo, consistent=True)
        dt = time.perf_counter() - t0
        diffs = []
        for q in range(6):
            den = float(np.abs(wo[q]).max())
            num = float(np.abs(wd[q].cpu().numpy() - wo[q]).max())
            diffs.append(num / den if den > 0 else num)
        ctx.close()
        out["cpu_baseline"] = {"value": m ** 3 * cpu_steps / dt, "unit": "cell*steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle IMEX run, {m}^3 cells (the input's grid), "
                                         f"{cpu_steps} steps, same options, {dt:.1f} s"}
        out["sample_max_rel_diff_vs_oracle"] = max(diffs)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chem-n", type=int, default=64 ** 3)
    ap.add_argument("--chem-steps", type=int, default=10)
    ap.add_argument("--chem-cpu-n", type=int, default=64 ** 3)
    ap.add_argument("--chem-cpu-reps", type=int, default=8)
    ap.add_argument("--no-chem", action="store_true")
    ap.add_argument("--no-vjp", action="store_true")
    ap.add_argument("--no-sod", action="store_true")
    ap.add_argument("--no-mri", action="store_true")
    ap.add_argument("--mri-n", type=int, default=128)
    ap.add_argument("--mri-steps", type=int, default=20)
    ap.add_argument("--imex-steps", type=int, default=20)
    ap.add_argument("--no-imex", action="store_true")
    ap.add_argument("--mri-table", choices=["ark437", "sdirk3"], default="ark437",
                    help="fast DIRK table of the multirate run (default: the reference "
                         "input's itable 112, ARK4(3)7L[2]SA's ESDIRK part)")
    ap.add_argument("--only", choices=["vjp", "sod", "mri", "chem", "fe10", "imex"], default=None,
                    help="run one extra workload alone and print its JSON (profiling runs)")
    ap.add_argument("--vjp-n", type=int, default=128)
    ap.add_argument("--vjp-reps", type=int, default=10)
    ap.add_argument("--profile-fe", action="store_true",
                    help="for hardware-counter runs: --steps fe() evaluations, then two "
                         "copies of a 1 GiB device buffer (the known-size calibration "
                         "kernel), and nothing else")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a multi-GPU fe() would "
              "exchange 3-deep halos between 3D blocks of the grid); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    api = _load("mvd_pyapi", os.path.join(PKG, "pyapi.py"))
    if args.only:
        problems = _load("mvd_problems", os.path.join(PKG, "problems.py"))
        quiet_gc()
        if args.only == "vjp":
            print(json.dumps(bench_vjp(api, problems, torch, args.vjp_n, 777, args.vjp_reps)))
        elif args.only == "mri":
            print(json.dumps(bench_mri(api, problems, torch, args.mri_n, args.mri_steps,
                                       cpu=not args.no_cpu, table=args.mri_table)))
        elif args.only == "imex":
            print(json.dumps(bench_imex(api, problems, torch, args.mri_n, args.imex_steps,
                                        cpu=not args.no_cpu)))
        elif args.only == "chem":
            print(json.dumps(bench_chem(api, problems, torch, args.chem_n, 4321,
                                        args.chem_steps)))
        elif args.only == "fe10":
            print(json.dumps(bench_fe_tracers(api, problems, torch, args.vjp_n, 777,
                                              args.vjp_reps)))
        else:
            print(json.dumps(bench_sod(api, problems, torch, cpu=not args.no_cpu)))
        return
    problems = _load("mvd_problems", os.path.join(PKG, "problems.py"))
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    ncell = n ** 3

    # inputs resident in HBM before the timed region
    w = [torch.from_numpy(a).to("cuda") for a in problems.blast3d(n, seed=args.seed)] + [None]
    wdot = [torch.empty_like(a) for a in w[:5]] + [None]
    ctx = api.Context(api.make_params(n, n, n, 0, (0,) * 6, 1.4))

    if args.profile_fe:
        for _ in range(args.steps):
            ctx.fe(w, wdot)
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"profile_fe": args.steps, "copy_bytes": a.numel() * 8}))
        ctx.close()
        return

    quiet_gc()
    for _ in range(args.warmup):
        ctx.fe(w, wdot)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.fe(w, wdot)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ms_per_step = (t1 - t0) / args.steps * 1e3
    gcells = ncell / (ms_per_step * 1e-3) / 1e9

    # plain comparator (reference operation order, one thread per cell)
    ctx.profile(False)
    ctx.fe(w, wdot, plain=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.plain_steps):
        ctx.fe(w, wdot, plain=True)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    plain_ms = (t3 - t2) / args.plain_steps * 1e3

    kernels = {}
    for name, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        avg = ms / cnt
        fl = kernel_flops(name, n)
        kernels[name] = {"avg_ms": avg, "launches": cnt,
                         "tflops": fl / (avg * 1e-3) / 1e12}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    dk = kernels[dom]
    fe_total_ms = sum(k["avg_ms"] for k in kernels.values())
    fe_flops = sum(kernel_flops(k, n) for k in kernels)
    roofline = {
        "bound": "fp64_valu",
        "kernel": dom,
        "achieved": dk["tflops"],
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": dk["tflops"] / PEAK_FP64_TFLOPS,
        "traffic": read_traffic(dom),
        "flops_per_launch": kernel_flops(dom, n),
        "avg_launch_ms": dk["avg_ms"],
    }
    extra = {
        "fe_kernels": kernels,
        "fe_kernel_sum_ms": fe_total_ms,
        "fe_fp64_tflops": fe_flops / (fe_total_ms * 1e-3) / 1e12,
        "fe_hbm_algorithmic_gbs": BYTES_PER_CELL_FE * ncell / (fe_total_ms * 1e-3) / 1e9,
        "fe_hbm_frac": BYTES_PER_CELL_FE * ncell / (fe_total_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
        "plain_gpu_gcells": ncell / (plain_ms * 1e-3) / 1e9,
        "speedup_vs_plain_gpu": plain_ms / ms_per_step,
    }

    if not args.no_chem:
        quiet_gc()
        extra["chem"] = bench_chem(api, problems, torch, args.chem_n, 4321, args.chem_steps)
    if not args.no_vjp:
        quiet_gc()
        extra["vjp"] = bench_vjp(api, problems, torch, args.vjp_n, 777, args.vjp_reps)
    if not args.no_sod:
        quiet_gc()
        extra["sod"] = bench_sod(api, problems, torch, cpu=not args.no_cpu)
    if not args.no_mri:
        quiet_gc()
        extra["mri"] = bench_mri(api, problems, torch, args.mri_n, args.mri_steps,
                                 cpu=not args.no_cpu, table=args.mri_table)
    if not args.no_imex:
        quiet_gc()
        extra["imex"] = bench_imex(api, problems, torch, args.mri_n, args.imex_steps,
                                   cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        v1, dt1 = cpu_baseline(args.cpu_n, 1)
        cpu = {"value": v1, "unit": "Gcell/s", "cores": 1, "kind": "port",
               "sample": f"one fe() of the blast IC at {args.cpu_n}^3 (periodic, nchem=0), "
                         f"oracle restatement of fEuler (serial, no profiler calls), -O2 "
                         f"without FMA contraction, {dt1:.1f} s"}
        nthr = min(16, os.cpu_count() or 1)
        vN, dtN = cpu_baseline(args.cpu_n, nthr)
        extra["cpu_openmp"] = {"value": vN, "unit": "Gcell/s", "cores": nthr, "seconds": dtN}
        extra["speedup_vs_cpu_1core"] = gcells / v1
        if not args.no_chem:
            c1, cdt = cpu_chem_baseline(args.chem_cpu_n, 4321, 1, args.chem_cpu_reps)
            extra["chem"]["cpu_baseline"] = {
                "value": c1, "unit": "cells/s", "cores": 1, "kind": "port",
                "sample": f"oracle SDIRK3 step of the {args.chem_cpu_n} config-3 cells, "
                          f"{args.chem_cpu_reps} times, {cdt:.1f} s"}
            cN, cdtN = cpu_chem_baseline(args.chem_cpu_n, 4321, nthr, 1)
            extra["chem"]["cpu_openmp"] = {"value": cN, "unit": "cells/s", "cores": nthr,
                                           "seconds": cdtN}

    line = {
        "metric": "fe() Gcell/s @256\u00b3 FP64; batched chem cells/s vs per-cell; "
                  "VJP/fwd cost ratio",
        "value": gcells,
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D fluid blast fe() only, {n}^3, periodic, nchem=0, "
                               f"gamma=1.4, seed {args.seed}",
                   "grid": [n, n, n], "nchem": 0},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

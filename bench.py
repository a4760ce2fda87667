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
# divergence per cell and field: 3 differences, 3 divides, 2 adds, 1 subtract
DIV_FLOPS_PER_VAR = 9
BYTES_PER_CELL_FE = 16 * 5          # read 5 fields + write 5 fields, FP64, nchem = 0
PEAK_FP64_TFLOPS = 78.6             # MI355X FP64 (vector and matrix) peak
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
    if name in ("fe_xface",):
        return faces * ff
    if name == "fe_march_y":
        return faces * ff + cells * nv * 6        # x and y terms of the divergence
    if name == "fe_march_z":
        return faces * ff + cells * nv * 3
    if name == "fe_yface":
        return faces * ff
    if name == "fe_zface_div":
        return faces * ff + cells * nv * DIV_FLOPS_PER_VAR
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-n", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the fe() halo exchange design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    api = _load("mvd_pyapi", os.path.join(PKG, "pyapi.py"))
    problems = _load("mvd_problems", os.path.join(PKG, "problems.py"))
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    ncell = n ** 3

    # inputs resident in HBM before the timed region
    w = [torch.from_numpy(a).to("cuda") for a in problems.blast3d(n, seed=args.seed)] + [None]
    wdot = [torch.empty_like(a) for a in w[:5]] + [None]
    ctx = api.Context(api.make_params(n, n, n, 0, (0,) * 6, 1.4))

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
        "bound": "mfma",
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

    cpu = None
    if not args.no_cpu:
        v1, dt1 = cpu_baseline(args.cpu_n, 1)
        cpu = {"value": v1, "unit": "Gcell/s", "cores": 1, "kind": "port",
               "sample": f"one fe() of the blast IC at {args.cpu_n}^3 (periodic, nchem=0), "
                         f"oracle restatement of fEuler, -O2 no FMA, {dt1:.1f} s"}
        nthr = min(16, os.cpu_count() or 1)
        vN, dtN = cpu_baseline(args.cpu_n, nthr)
        extra["cpu_openmp"] = {"value": vN, "unit": "Gcell/s", "cores": nthr, "seconds": dtN}
        extra["speedup_vs_cpu_1core"] = gcells / v1

    line = {
        "metric": "fe() Gcell/s @256^3 FP64",
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

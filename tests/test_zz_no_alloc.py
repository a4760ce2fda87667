"""The allocation-free batched chemistry step (mvd_chem_dirk_step_noalloc) through the C ABI
against the allocating entry point (mvd_chem_dirk_step, plain = 0): bitwise equal states,
temperatures and statistics on consecutive calls of decreasing size (the shared scratch
array keeps the last call's contents: sorted 50000 cells, then sorted 20000, then 3000
cells in natural order, then 20000 again on 16 waves so the work queue runs), no change of
the device's free memory across those calls, and the size limit rejected before any work.
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

pytestmark = pytest.mark.gpu
MAX_CELLS = 1 << 20


def make(mvd, torch, n, seed):
    y, _ = problems.chem_cells(n, seed=seed)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    return ctx, yd


def test_noalloc_step_matches_allocating_step(mvd, torch_gpu):
    torch = torch_gpu
    cases = [(50000, 41, 0), (20000, 42, 0), (3000, 43, 0), (20000, 44, 16)]
    pairs = []
    for n, seed, waves in cases:
        print(f"[noalloc] setup n={n} seed={seed} waves={waves}", flush=True)
        ref = make(mvd, torch, n, seed)
        na = make(mvd, torch, n, seed)
        if waves:
            ref[0].dirk_set_shape(256, waves)
            na[0].dirk_set_shape(256, waves)
        pairs.append((ref, na))
    # the allocating path first (its work buffers are allocated here)
    ref_out = []
    for (ctx, yd), _ in pairs:
        f, st = ctx.dirk_step(yd, 1e3)
        ref_out.append((f, st, ctx.chem_state("Ts", torch).clone()))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    na_steps = []
    for (_, (ctx, yd)), (n, seed, waves) in zip(pairs, cases):
        print(f"[noalloc] step n={n} waves={waves}", flush=True)
        na_steps.append(ctx.dirk_step_noalloc(yd, 1e3))
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0             # (no tensor is created between the two readings)
    na_out = [(f, st, ctx.chem_state("Ts", torch).clone())
              for (f, st), (_, (ctx, _)) in zip(na_steps, pairs)]
    for ((rctx, ry), (nctx, ny)), (rf, rst, rT), (nf, nst, nT) in zip(pairs, ref_out, na_out):
        assert rf == nf == 0
        assert torch.equal(ry, ny)
        assert torch.equal(rT, nT)
        for k in ("newton_iters", "max_cell_iters", "failed_cells", "setups", "hist"):
            assert rst[k] == nst[k], k
        rctx.close()
        nctx.close()


def test_noalloc_step_rejects_too_many_cells(mvd, torch_gpu):
    torch = torch_gpu
    n = MAX_CELLS + 1
    print(f"[noalloc] limit n={n}", flush=True)
    ctx, yd = make(mvd, torch, n, 45)
    before = yd.clone()
    with pytest.raises(mvd.MvdError) as e:
        ctx.dirk_step_noalloc(yd, 1e3)
    assert e.value.code == mvd.MVD_BAD_ARGUMENT
    assert torch.equal(yd, before)
    ctx.close()

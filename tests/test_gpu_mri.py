"""Multirate evolution on the GPU (mvd_mri_evolve) and the forced DIRK step
(mvd_chem_dirk_step_forced) through the C ABI, against the CPU oracle (pinned in
tests/test_oracle_mri.py and, for its pieces, test_oracle_fe.py / test_oracle_chem.py /
test_oracle_dirk.py).

Tolerances: the forced step 1e-10 per species (relative to that species' largest magnitude,
as the unforced DIRK tests); the multirate run 1e-8 on every field and species after three
slow steps (the integrated-state tolerance); graph replay bit for bit.
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


def test_forced_dirk_matches_oracle(mvd, torch_gpu):
    torch = torch_gpu
    n = 4096
    print(f"[forced dirk] n={n}", flush=True)
    y, _ = problems.chem_cells(n, seed=61)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    rng = np.random.default_rng(62)
    force = rng.standard_normal(10 * n) * 1e-5
    assert orc.dirk_step_forced(oc, orc.dirk_opts(), 1e3, yo, force) == 0
    fd = torch.from_numpy(force).to("cuda")
    assert ctx.dirk_step_forced(yd, 1e3, fd) == 0
    got = yd.cpu().numpy().reshape(n, 10)
    ref = yo.reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= 1e-10, s
    ctx.close()


def mri_pair(mvd, torch, dims, bc):
    n = dims[0] * dims[1] * dims[2]
    fluid, chem = problems.primordial_blast_mri(*dims)
    ctx = mvd.Context(mvd.make_params(*dims, 10, bc, gamma=5 / 3, units=UNITS))
    ctx.chem_setup(orc.load_tables(), n)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
         [torch.from_numpy(chem.copy()).to("cuda")]
    ctx.chem_prepare(wd[5])
    oc = orc.Chem(n)
    y = chem.copy()
    oc.prepare(y)
    oc.extra()
    wo = [a.copy() for a in fluid] + [y]
    return ctx, wd, oc, wo


@pytest.mark.parametrize("consistent", [False, True])
def test_mri_matches_oracle(mvd, torch_gpu, consistent):
    torch = torch_gpu
    dims, bc = (12, 10, 8), (3,) * 6
    n = 12 * 10 * 8
    print(f"[mri] dims={dims} 3 slow steps consistent={consistent}", flush=True)
    ctx, wd, oc, wo = mri_pair(mvd, torch, dims, bc)
    g = orc.make_grid(*dims, 10, bc, gamma=5 / 3)
    assert orc.mri_evolve(g, oc, EU, orc.dirk_opts(rhs_scale=UNITS[2]), 0.05, 3, 0.01, wo,
                          consistent) == 0
    fast = mvd.sdirk3(rhs_scale=UNITS[2])
    assert ctx.mri_evolve(wd, 0.05, 3, 0.01, fast, consistent) == 0
    for q in range(5):
        assert normwise_err(wd[q].cpu().numpy(), wo[q]) <= 1e-8, q
    got = wd[5].cpu().numpy().reshape(n, 10)
    ref = wo[5].reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= 1e-8, ("chem", s)
    ctx.close()


def test_mri_graph_replay_is_exact(mvd, torch_gpu):
    """3 slow steps in one call equal 1 + 2 calls bit for bit (eager step, captured step
    replayed), also with a plain-path chemistry step between the calls, and a changed fast
    bound recaptures."""
    torch = torch_gpu
    dims, bc = (10, 8, 8), (3,) * 6
    fast = mvd.sdirk3(rhs_scale=UNITS[2])
    ctx, a, _, _ = mri_pair(mvd, torch, dims, bc)
    Ts0 = ctx.chem_state("Ts", torch).clone()
    a0 = [t.clone() for t in a]
    ctx.mri_evolve(a, 0.05, 3, 0.01, fast)
    ta = ctx.chem_state("Ts", torch).clone()
    b = [t.clone() for t in a0]
    ctx.chem_state_set("Ts", Ts0)
    ctx.mri_evolve(b, 0.05, 1, 0.01, fast)
    # a plain-path step in between takes a larger chemistry work buffer, which the captured
    # slow step must not keep using (it is captured again)
    Ts1 = ctx.chem_state("Ts", torch).clone()
    dT1 = ctx.chem_state("dTs_ge", torch).clone()
    scratch = b[5].clone()
    ctx.dirk_step(scratch, 1e3, plain=True)
    ctx.chem_state_set("Ts", Ts1)
    ctx.chem_state_set("dTs_ge", dT1)
    ctx.mri_evolve(b, 0.05, 2, 0.01, fast)
    for q in range(6):
        assert torch.equal(a[q], b[q]), q
    assert torch.equal(ta, ctx.chem_state("Ts", torch))
    c = [t.clone() for t in a0]
    ctx.chem_state_set("Ts", Ts0)
    assert ctx.mri_evolve(c, 0.05, 2, 0.02, fast) == 0
    assert not torch.equal(c[5], a[5])
    ctx.close()


def test_step_rejection_matches_oracle(mvd, torch_gpu):
    """Per-cell step rejection (max_halvings): with a Newton budget too small for the full
    step many cells fail; retried in halves most converge, the rest keep their last accepted
    state. Device and oracle take the same substeps (states to 1e-10, the same failed
    cells)."""
    torch = torch_gpu
    n = 4096
    print(f"[step rejection] n={n}", flush=True)
    y, _ = problems.chem_cells(n, seed=71)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    y0 = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(y0)
    Ts0 = ctx.chem_state("Ts", torch).clone()
    results = {}
    for mh in (0, 6):
        oc = orc.Chem(n)
        yo = y.copy()
        oc.prepare(yo)
        oc.extra()
        opts_o = orc.dirk_opts(max_iters=2, max_setups=0, max_halvings=mh)
        fo, _ = orc.dirk_step(oc, opts_o, 1e5, yo)
        yd = y0.clone()
        ctx.chem_state_set("Ts", Ts0)
        fd, _ = ctx.dirk_step(yd, 1e5, mvd.sdirk3(max_iters=2, max_setups=0, max_halvings=mh))
        results[mh] = (fo, fd)
        assert fd == fo
        if mh:
            got = yd.cpu().numpy().reshape(n, 10)
            ref = yo.reshape(n, 10)
            for s in range(10):
                assert normwise_err(got[:, s], ref[:, s]) <= 1e-10, s
            Tg = ctx.chem_state("Ts", torch).cpu().numpy()
            assert np.max(np.abs(Tg - oc.array("Ts")) / oc.array("Ts")) <= 1e-10
    # h = 1e5 s: most cells fail without halving, a few even at the finest size
    assert results[0][0] > 500 and 0 < results[6][0] < results[0][0]
    ctx.close()


def test_mri_reports_illegal_state(mvd, torch_gpu):
    """The slow RHS (fslow, tracer kernels) meeting a cell of negative density makes the
    multirate run return MVD_ILLEGAL_STATE; mvd_last_state_flags is not changed by it."""
    torch = torch_gpu
    dims, bc = (8, 8, 6), (3,) * 6
    ctx, wd, _, _ = mri_pair(mvd, torch, dims, bc)
    assert ctx.last_state_flags() == 0
    c = dims[0] * dims[1] * 2 + 3
    wd[0][c] = -abs(float(wd[0][c]))
    fast = mvd.sdirk3(rhs_scale=UNITS[2])
    with pytest.raises(mvd.MvdError) as e:
        ctx.mri_evolve(wd, 0.05, 1, 0.01, fast)
    assert e.value.code == mvd.MVD_ILLEGAL_STATE
    assert ctx.last_state_flags() == 0
    ctx.close()


def test_mri_graph_recaptured_after_chemistry_setup(mvd, torch_gpu):
    """A captured slow step holds the chemistry arrays' addresses and the redshift by value.
    After mvd_chem_setup (new arrays, another redshift) and a plain DIRK step (which may
    reallocate the work buffer), the next multirate run must capture again: its result
    equals the same run in a fresh context bit for bit."""
    torch = torch_gpu
    dims, bc = (10, 8, 8), (3,) * 6
    n = dims[0] * dims[1] * dims[2]
    fast = mvd.sdirk3(rhs_scale=UNITS[2])
    fluid, chem = problems.primordial_blast_mri(*dims)
    ctx, a, _, _ = mri_pair(mvd, torch, dims, bc)
    assert ctx.mri_evolve(a, 0.05, 2, 0.01, fast) == 0        # eager step + captured step
    # new chemistry setup with another redshift, re-prepared from the initial species
    ctx.chem_setup(orc.load_tables(), n, current_z=0.5)
    b = [torch.from_numpy(x.copy()).to("cuda") for x in fluid] + \
        [torch.from_numpy(chem.copy()).to("cuda")]
    ctx.chem_prepare(b[5])
    Ts0 = ctx.chem_state("Ts", torch).clone()
    dT0 = ctx.chem_state("dTs_ge", torch).clone()
    scratch = b[5].clone()
    ctx.dirk_step(scratch, 1e3, plain=True)
    ctx.chem_state_set("Ts", Ts0)
    ctx.chem_state_set("dTs_ge", dT0)
    assert ctx.mri_evolve(b, 0.05, 3, 0.01, fast) == 0
    fresh = mvd.Context(mvd.make_params(*dims, 10, bc, gamma=5 / 3, units=UNITS))
    fresh.chem_setup(orc.load_tables(), n, current_z=0.5)
    c = [torch.from_numpy(x.copy()).to("cuda") for x in fluid] + \
        [torch.from_numpy(chem.copy()).to("cuda")]
    fresh.chem_prepare(c[5])
    assert fresh.mri_evolve(c, 0.05, 3, 0.01, fast) == 0
    for q in range(6):
        assert torch.equal(b[q], c[q]), q
    assert torch.equal(ctx.chem_state("Ts", torch), fresh.chem_state("Ts", torch))
    ctx.close()
    fresh.close()

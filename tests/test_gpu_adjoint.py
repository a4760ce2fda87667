"""Checkpointed discrete adjoint of the chemistry DIRK steps on the GPU
(mvd_chem_dirk_adjoint) through the C ABI, against the CPU oracle's adjoint (no
checkpointing; pinned by invariants, central differences and a dot-product test in
tests/test_oracle_adjoint.py).

Tolerances: final states 1e-10 and adjoints 1e-8 per component, relative to that
component's largest magnitude over the batch (device log/exp/pow differ from glibc by an
ulp and the stiff steps carry it); invariant cotangents pass through to 1e-12; results do
not depend on the checkpoint interval (bitwise: recomputation repeats the forward run).
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


def pair(mvd, torch, n, seed):
    y, _ = problems.chem_cells(n, seed=seed)
    oc = orc.Chem(n)
    yo = y.copy()
    oc.prepare(yo)
    oc.extra()
    rng = np.random.default_rng(seed)
    yo *= 1.0 + 0.3 * rng.uniform(-1, 1, yo.size)
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 10))
    ctx.chem_setup(orc.load_tables(), n)
    yd = torch.from_numpy(y.copy()).to("cuda")
    ctx.chem_prepare(yd)
    yd.copy_(torch.from_numpy(yo))
    # the same temperature warm start on both sides
    ctx.chem_state_set("Ts", torch.from_numpy(oc.array("Ts").copy()).to("cuda"))
    return oc, yo, ctx, yd


def energy_cotangent(oc, n):
    """d(sum_cells mdensity * ge) / d y_N in scaled variables: the domain-integrated gas
    energy density (cell volume 1)."""
    lam = np.zeros((n, 10))
    lam[:, 9] = oc.array("mdensity") * oc.array("scale").reshape(n, 10)[:, 9]
    return lam.reshape(-1)


# 8192 cells: above the batched step's ordering threshold, so the adjoint runs its cells in
# temperature order (thread u takes the u-th cell of the order)
@pytest.mark.parametrize("n,nsteps,K,h", [(512, 12, 5, 200.0), (1024, 50, 10, 50.0),
                                          (8192, 12, 5, 200.0)])
def test_adjoint_matches_oracle(mvd, torch_gpu, n, nsteps, K, h):
    torch = torch_gpu
    print(f"[chem adjoint] n={n} steps={nsteps} ckpt={K} h={h}", flush=True)
    oc, yo, ctx, yd = pair(mvd, torch, n, 41)
    rng = np.random.default_rng(5)
    lamT = energy_cotangent(oc, n) * (1.0 + 0.1 * rng.standard_normal(10 * n))
    yref = yo.copy()
    failed, lref = orc.dirk_adjoint(oc, orc.dirk_opts(), h, nsteps, yref, lamT)
    assert failed == 0
    lt = torch.from_numpy(lamT).to("cuda")
    l0 = torch.empty_like(lt)
    assert ctx.dirk_adjoint(yd, h, nsteps, K, lt, l0) == 0
    yg = yd.cpu().numpy().reshape(n, 10)
    lg = l0.cpu().numpy().reshape(n, 10)
    for s in range(10):
        assert normwise_err(yg[:, s], yref.reshape(n, 10)[:, s]) <= 1e-10, ("y", s)
        assert normwise_err(lg[:, s], lref.reshape(n, 10)[:, s]) <= 1e-8, ("lam", s)
    ctx.close()


def test_adjoint_invariants_and_checkpoint_interval(mvd, torch_gpu):
    torch = torch_gpu
    n, nsteps, h = 2048, 9, 200.0
    print(f"[chem adjoint invariants] n={n}", flush=True)
    oc, yo, ctx, yd0 = pair(mvd, torch, n, 43)
    Ts0 = ctx.chem_state("Ts", torch).clone()
    sc = ctx.chem_state("scale", torch).cpu().numpy().reshape(n, 10)
    for c in ([2, 2, 1, 1, 1, 0, 0, 0, 0, 0.0], [0, 0, 0, 0, 0, 1, 1, 1, 0, 0.0],
              [0, 1, 0, 1, -1, 0, 1, 2, -1, 0.0]):
        lamT = (np.asarray(c)[None, :] * sc).reshape(-1)
        yd = yd0.clone()
        ctx.chem_state_set("Ts", Ts0)
        lt = torch.from_numpy(lamT).to("cuda")
        l0 = torch.empty_like(lt)
        ctx.dirk_adjoint(yd, h, nsteps, 4, lt, l0)
        err = np.abs(l0.cpu().numpy() - lamT).reshape(n, 10).max(axis=1)
        assert np.all(err <= 1e-12 * np.abs(lamT).reshape(n, 10).max(axis=1))
    lamT = torch.from_numpy(energy_cotangent(oc, n)).to("cuda")
    outs = []
    for K in (1, 4, nsteps):
        yd = yd0.clone()
        ctx.chem_state_set("Ts", Ts0)
        l0 = torch.empty_like(lamT)
        ctx.dirk_adjoint(yd, h, nsteps, K, lamT, l0)
        outs.append((yd.clone(), l0))
    for yk, lk in outs[1:]:
        assert torch.equal(yk, outs[0][0]) and torch.equal(lk, outs[0][1])
    ctx.close()

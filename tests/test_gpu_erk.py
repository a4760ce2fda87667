"""Fixed-step explicit evolution of fe() (mvd_erk_evolve) and the N_Vector operations on
the GPU through the C ABI, against the CPU oracle (itself pinned in test_oracle_erk.py).

Tolerance: 1e-8 relative (per field, to its largest magnitude) on the 2000-step Sod run of
configuration 1 -- the integrated-state tolerance; 1e-10 on short 3D runs. The device and
the oracle evaluate the same stage algebra; they differ by the ulp-level differences of
fe() (contracted products in the fast kernels) that the steps carry forward.
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


def to_dev(torch, w):
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1).copy())
            .to("cuda") for a in w]


def oracle_evolve(g, w0, h, nsteps):
    w = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1).copy() for a in w0]
    assert orc.erk_evolve(g, orc.erk(*orc.bs3_table()), h, nsteps, w) == 0
    return w


def test_sod_config1_matches_oracle(mvd, torch_gpu):
    torch = torch_gpu
    nx = 128
    print("[erk sod] 128x3x3, 2000 steps of 1e-4", flush=True)
    w0 = [a for a in problems.sod(nx)[:5]] + [np.zeros(1)]
    ref = oracle_evolve(orc.make_grid(nx, 3, 3, 0, (1,) * 6), w0, 1e-4, 2000)
    ctx = mvd.Context(mvd.make_params(nx, 3, 3, 0, (1,) * 6))
    wd = to_dev(torch, w0[:5]) + [None]
    assert ctx.erk_evolve(wd, 1e-4, 2000) == 0
    for q in range(5):
        assert normwise_err(wd[q].cpu().numpy(), ref[q]) <= 1e-8, q
    ctx.close()


@pytest.mark.parametrize("dims,nchem,bc", [((20, 18, 16), 2, (0,) * 6),
                                           ((16, 12, 10), 0, (3, 1, 1, 3, 0, 0))])
def test_erk_3d_matches_oracle(mvd, torch_gpu, dims, nchem, bc):
    torch = torch_gpu
    print(f"[erk 3d] dims={dims} nchem={nchem} bc={bc}", flush=True)
    w0 = problems.smooth_state(*dims, nchem, seed=9)
    ref = oracle_evolve(orc.make_grid(*dims, nchem, bc), w0, 2e-3, 12)
    ctx = mvd.Context(mvd.make_params(*dims, nchem, bc))
    wd = to_dev(torch, w0)
    if nchem == 0:
        wd[5] = None
    assert ctx.erk_evolve(wd, 2e-3, 12) == 0
    for q in range(5 + (nchem > 0)):
        assert normwise_err(wd[q].cpu().numpy(), ref[q]) <= 1e-10, q
    ctx.close()


def test_erk_graph_replay_and_recapture(mvd, torch_gpu):
    """10 + 10 steps equal 20 steps bitwise (one captured step replayed); a new step size
    recaptures; an illegal state is reported, not hidden."""
    torch = torch_gpu
    dims = (24, 8, 8)
    w0 = problems.smooth_state(*dims, 0, seed=4)
    ctx = mvd.Context(mvd.make_params(*dims, 0, (0,) * 6))
    a = to_dev(torch, w0[:5]) + [None]
    b = to_dev(torch, w0[:5]) + [None]
    ctx.erk_evolve(a, 1e-3, 10)
    ctx.erk_evolve(a, 1e-3, 10)
    ctx.erk_evolve(b, 1e-3, 20)
    for q in range(5):
        assert torch.equal(a[q], b[q]), q
    c = to_dev(torch, w0[:5]) + [None]
    ctx.erk_evolve(c, 2e-3, 10)
    ref = oracle_evolve(orc.make_grid(*dims, 0, (0,) * 6), w0, 2e-3, 10)
    for q in range(5):
        assert normwise_err(c[q].cpu().numpy(), ref[q]) <= 1e-10, q
    bad = to_dev(torch, w0[:5]) + [None]
    bad[4][7] = -1.0                              # negative total energy in one cell
    assert ctx.erk_evolve(bad, 1e-3, 2, allow_illegal=True) == mvd.MVD_ILLEGAL_STATE
    ctx.close()


def test_nvec_linear_combination_and_wrms(mvd, torch_gpu):
    torch = torch_gpu
    ctx = mvd.Context(mvd.make_params(4, 4, 4, 0))
    rng = np.random.default_rng(3)
    n = 100003
    X = [rng.standard_normal(n) for _ in range(19)]
    c = rng.standard_normal(19)
    Xd = [torch.from_numpy(x).to("cuda") for x in X]
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.linear_combination(c[:5], Xd[:5], z)
    ref = c[0] * X[0]
    for k in range(1, 5):
        ref = ref + c[k] * X[k]
    assert np.array_equal(z.cpu().numpy(), ref)           # same order, no contraction
    ctx.linear_combination(c, Xd, z)                      # more than one launch (> 16)
    ref = c[0] * X[0]
    for k in range(1, 19):
        ref = ref + c[k] * X[k]
    assert np.array_equal(z.cpu().numpy(), ref)
    x0 = Xd[0].clone()
    ctx.linear_combination([2.0, -1.0], [x0, Xd[1]], x0)  # z = X[0] in place
    assert np.array_equal(x0.cpu().numpy(), 2.0 * X[0] - X[1])
    lens = [n, 777, 5000]
    xs = [rng.standard_normal(m) for m in lens]
    ws = [rng.uniform(0.5, 2.0, m) for m in lens]
    got = ctx.wrms_norm([torch.from_numpy(a).to("cuda") for a in xs],
                        [torch.from_numpy(a).to("cuda") for a in ws])
    want = np.sqrt(sum(((a * b) ** 2).sum() for a, b in zip(xs, ws)) / sum(lens))
    assert abs(got - want) <= 1e-14 * want
    ctx.close()

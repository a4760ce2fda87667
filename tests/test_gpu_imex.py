"""IMEX additive Runge-Kutta evolution on the GPU (mvd_imex_evolve) through the C ABI
against the CPU oracle (pinned in tests/test_oracle_imex.py by the table's order conditions
and third-order convergence of its explicit and implicit halves).

Tolerance: 1e-8 on every field and species after three steps (the integrated-state
tolerance); the table is compared bit for bit.
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


def test_ark324_table_matches_oracle(mvd, torch_gpu):
    d, o = mvd.ark324(), orc.ark324_table()
    assert d.stages == o.stages == 4
    for f in ("Ae", "Ai", "b", "c"):
        assert list(getattr(d, f)) == list(getattr(o, f)), f


@pytest.mark.parametrize("dims,consistent,at_rest,pair", [((10, 8, 8), False, False, "ark324"),
                                                          ((16, 16, 16), True, False, "ark324"),
                                                          ((8, 8, 8), False, True, "ark324"),
                                                          ((10, 8, 8), False, False, "ark437"),
                                                          ((16, 16, 16), True, False, "ark437")])
def test_imex_matches_oracle(mvd, torch_gpu, dims, consistent, at_rest, pair):
    """ARK3(2)4L[2]SA and the reference IMEX inputs' ARK4(3)7L[2]SA (etable 13 / itable 112,
    explicit first stage) against the oracle's driver."""
    torch = torch_gpu
    n = dims[0] * dims[1] * dims[2]
    print(f"[imex] dims={dims} consistent={consistent} at_rest={at_rest} pair={pair}",
          flush=True)
    tab_o = orc.ark437_pair() if pair == "ark437" else orc.ark324_table()
    tab_d = mvd.ark437_pair() if pair == "ark437" else mvd.ark324()
    fluid, chem = problems.primordial_blast_mri(*dims, bump=3.0, width=0.25)
    if at_rest:
        c0 = n // 2 + dims[0] // 2
        fluid = [np.full(n, a[c0]) for a in fluid]
        chem = np.tile(chem.reshape(n, 10)[c0], n)
    bc = (0,) * 6
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
    g = orc.make_grid(*dims, 10, bc, gamma=5 / 3)
    h, nsteps = 1e-4, 3
    assert orc.imex_evolve(g, oc, EU, tab_o, orc.dirk_opts(rhs_scale=UNITS[2]),
                           h, nsteps, wo, consistent) == 0
    assert ctx.imex_evolve(wd, h, nsteps, mvd.sdirk3(rhs_scale=UNITS[2]), table=tab_d,
                           consistent=consistent) == 0
    for q in range(5):
        if np.max(np.abs(wo[q])) > 0:
            assert normwise_err(wd[q].cpu().numpy(), wo[q]) <= 1e-8, q
        else:
            assert np.max(np.abs(wd[q].cpu().numpy())) == 0, q
    got = wd[5].cpu().numpy().reshape(n, 10)
    ref = wo[5].reshape(n, 10)
    for s in range(10):
        assert normwise_err(got[:, s], ref[:, s]) <= 1e-8, ("chem", s)
    ctx.close()


def test_imex_rejects_bad_tables(mvd, torch_gpu):
    torch = torch_gpu
    dims = (4, 4, 4)
    ctx = mvd.Context(mvd.make_params(*dims, 10, (0,) * 6, gamma=5 / 3, units=UNITS))
    ctx.chem_setup(orc.load_tables(), 64)
    fluid, chem = problems.primordial_blast_mri(*dims)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
         [torch.from_numpy(chem.copy()).to("cuda")]
    ctx.chem_prepare(wd[5])
    t = mvd.ark324()
    t.Ae[1] = 0.5                          # above the diagonal
    with pytest.raises(mvd.MvdError):
        ctx.imex_evolve(wd, 1e-4, 1, mvd.sdirk3(rhs_scale=UNITS[2]), table=t)
    t = mvd.ark324()
    t.Ai[9] = -0.1                         # negative diagonal
    with pytest.raises(mvd.MvdError):
        ctx.imex_evolve(wd, 1e-4, 1, mvd.sdirk3(rhs_scale=UNITS[2]), table=t)
    ctx.close()


def test_imex_reports_illegal_state(mvd, torch_gpu):
    """fexpl's fe() runs the tracer kernels; a cell with negative density makes the run
    return MVD_ILLEGAL_STATE (the stage evaluations OR their legal-state bits into the run's
    flag), while mvd_last_state_flags keeps what the last standalone fe() call left."""
    torch = torch_gpu
    dims = (8, 6, 6)
    n = dims[0] * dims[1] * dims[2]
    ctx = mvd.Context(mvd.make_params(*dims, 10, (0,) * 6, gamma=5 / 3, units=UNITS))
    ctx.chem_setup(orc.load_tables(), n)
    fluid, chem = problems.primordial_blast_mri(*dims, bump=3.0, width=0.25)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in fluid] + \
         [torch.from_numpy(chem.copy()).to("cuda")]
    ctx.chem_prepare(wd[5])
    assert ctx.last_state_flags() == 0
    # (fexpl rebuilds the total energy from the gas energy, so the bad value is a density)
    wd[0][n // 3] = -abs(float(wd[0][n // 3]))
    with pytest.raises(mvd.MvdError) as e:
        ctx.imex_evolve(wd, 1e-4, 1, mvd.sdirk3(rhs_scale=UNITS[2]))
    assert e.value.code == mvd.MVD_ILLEGAL_STATE
    assert ctx.last_state_flags() == 0
    ctx.close()

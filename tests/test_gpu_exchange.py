"""Ghost exchange on the GPU (mvd_exchange, mvd_pack1d) through the C ABI against the oracle
(pinned in test_oracle_exchange.py by the reference's index-coded exchange test and its
boundary fills). Copies and sign flips only, so the bar is bitwise equality; the index-coded
known-answer test of communication_test_main.cpp is also run on the device directly."""
import itertools

import numpy as np
import pytest

import orc
from test_oracle_exchange import buffer_view, code, coded_state

pytestmark = pytest.mark.gpu

CASES = [((9, 7, 6), 0, (0,) * 6), ((8, 6, 7), 2, (1,) * 6), ((6, 9, 5), 10, (3,) * 6),
         ((7, 5, 8), 4, (2,) * 6), ((6, 7, 5), 1, (1, 2, 0, 0, 2, 3))]


def dev_state(torch, w, nchem):
    return [torch.from_numpy(a.copy()).to("cuda") for a in w[:5]] + \
           ([torch.from_numpy(w[5].copy()).to("cuda")] if nchem else [None])


@pytest.mark.parametrize("dims,nchem,bc", CASES)
def test_exchange_matches_oracle(mvd, torch_gpu, dims, nchem, bc):
    torch = torch_gpu
    print(f"[exchange] dims={dims} nchem={nchem} bc={bc}", flush=True)
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims) + nchem)
    n = nx * ny * nz
    w = [rng.standard_normal(n) for _ in range(5)] + \
        [rng.standard_normal(n * nchem) if nchem else np.zeros(1)]
    ref = orc.exchange(orc.make_grid(nx, ny, nz, nchem, bc), w)
    ctx = mvd.Context(mvd.make_params(nx, ny, nz, nchem, bc))
    wd = dev_state(torch, w, nchem)
    recv = [torch.full((r.size,), float("nan"), dtype=torch.float64, device="cuda") for r in ref]
    ctx.exchange(wd, recv)
    for f in range(6):
        assert np.array_equal(recv[f].cpu().numpy(), ref[f]), f
    # one face only (the others skipped)
    only = [None] * 6
    only[3] = torch.zeros_like(recv[3])
    ctx.exchange(wd, only)
    assert np.array_equal(only[3].cpu().numpy(), ref[3])
    # stencil packs of every boundary-touching cell and some interior ones, each direction
    cells = [(i, j, k) for i, j, k in itertools.product(range(nx), range(ny), range(nz))
             if min(i, j, k) < 3 or i >= nx - 3 or j >= ny - 3 or k >= nz - 3]
    ijk = torch.tensor(np.array(cells, dtype=np.int32)).to("cuda")
    g = orc.make_grid(nx, ny, nz, nchem, bc)
    nv = 5 + nchem
    for d in range(3):
        out = torch.empty(len(cells) * 6 * nv, dtype=torch.float64, device="cuda")
        ctx.pack1d(wd, d, ijk, out)
        got = out.cpu().numpy().reshape(len(cells), 6, nv)
        for q, (i, j, k) in enumerate(cells):
            assert np.array_equal(got[q], orc.gather(g, w, d, i, j, k)), (d, i, j, k)
    bad = torch.tensor(np.array([[0, 0, 0], [nx, 0, 0]], dtype=np.int32)).to("cuda")
    with pytest.raises(mvd.MvdError):
        ctx.pack1d(wd, 0, bad, torch.empty(2 * 6 * nv, dtype=torch.float64, device="cuda"))
    ctx.close()


def test_exchange_index_coded_kat_on_device(mvd, torch_gpu):
    """communication_test_main.cpp:133-605 on the device: periodic boundaries, one task,
    values 0.TTsXXXyyyZZZ; every receive slot holds the coded value of its source cell."""
    torch = torch_gpu
    nx, ny, nz, nchem = 8, 6, 7, 4
    nv = 5 + nchem
    w = coded_state(nx, ny, nz, nchem)
    ctx = mvd.Context(mvd.make_params(nx, ny, nz, nchem, (0,) * 6))
    sizes = [3 * ny * nz] * 2 + [3 * nx * nz] * 2 + [3 * nx * ny] * 2
    recv = [torch.empty(m * nv, dtype=torch.float64, device="cuda") for m in sizes]
    ctx.exchange(dev_state(torch, w, nchem), recv)
    W, E, S, N, B, F = [r.cpu().numpy() for r in recv]
    Wv, Ev = buffer_view(W, nv, ny, nz), buffer_view(E, nv, ny, nz)
    Sv, Nv = buffer_view(S, nv, nx, nz), buffer_view(N, nv, nx, nz)
    Bv, Fv = buffer_view(B, nv, nx, ny), buffer_view(F, nv, nx, ny)
    tol, errs = 1e-12, 0
    for k, j, i, v in itertools.product(range(nz), range(ny), range(3), range(nv)):
        errs += abs(Wv[k, j, i, v] - code(v, i + nx - 3, j, k)) > tol
        errs += abs(Ev[k, j, i, v] - code(v, i, j, k)) > tol
    for k, j, i, v in itertools.product(range(nz), range(3), range(nx), range(nv)):
        errs += abs(Sv[k, i, j, v] - code(v, i, j + ny - 3, k)) > tol
        errs += abs(Nv[k, i, j, v] - code(v, i, j, k)) > tol
    for k, j, i, v in itertools.product(range(3), range(ny), range(nx), range(nv)):
        errs += abs(Bv[j, i, k, v] - code(v, i, j, k + nz - 3)) > tol
        errs += abs(Fv[j, i, k, v] - code(v, i, j, k)) > tol
    assert errs == 0
    ctx.close()

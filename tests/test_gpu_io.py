"""Solution output / restart input from and to device arrays (libmvd_io.so through the C
ABI): the file written from HBM is the one written from the same values on the host, and
reading it back into device arrays reproduces the state (the reference's I/O test,
io_test_main.cpp:114-227, at unit factors 1 exactly)."""
import numpy as np
import pytest

from test_io import coded_state, params

pytestmark = pytest.mark.gpu


def test_device_round_trip_matches_host_file(mvd, torch_gpu, tmp_path):
    torch = torch_gpu
    print("[io] device round trip 7x5x4 nchem=10", flush=True)
    p = params(mvd, units=(3.0e70, 3.0857e30, 1.0e11))
    w = coded_state(7, 5, 4, 10)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in w]
    (tmp_path / "dev").mkdir()
    (tmp_path / "host").mkdir()
    mvd.output_solution(tmp_path / "dev", p, 8.5, 1e-3, 5, wd, mvd.RunOptions())
    mvd.output_solution(tmp_path / "host", p, 8.5, 1e-3, 5, w, mvd.RunOptions())
    # device arrays untouched by the unit conversion
    for q in range(6):
        assert np.array_equal(wd[q].cpu().numpy(), w[q])
    hd = [np.empty_like(a) for a in w]
    hh = [np.empty_like(a) for a in w]
    assert mvd.read_restart(tmp_path / "dev", p, 5, hd) == 8.5
    assert mvd.read_restart(tmp_path / "host", p, 5, hh) == 8.5
    for q in range(6):
        assert np.array_equal(hd[q], hh[q]), q
    rd = [torch.full_like(t, float("nan")) for t in wd]
    assert mvd.read_restart(tmp_path / "dev", p, 5, rd) == 8.5
    for q in range(6):
        assert np.array_equal(rd[q].cpu().numpy(), hh[q]), q
    assert ((tmp_path / "dev" / "restart_parameters.txt").read_text() ==
            (tmp_path / "host" / "restart_parameters.txt").read_text())


def test_device_round_trip_exact_at_unit_factors_one(mvd, torch_gpu, tmp_path):
    torch = torch_gpu
    print("[io] device exact round trip 16x9x5 nchem=10", flush=True)
    p = params(mvd, nx=16, ny=9, nz=5)
    w = coded_state(16, 9, 5, 10)
    wd = [torch.from_numpy(a.copy()).to("cuda") for a in w]
    mvd.output_solution(tmp_path, p, 8.5, 0.0, 5, wd)
    rd = [torch.zeros_like(t) for t in wd]
    assert abs(mvd.read_restart(tmp_path, p, 5, rd) - 8.5) <= 1e-14
    for q in range(6):
        assert torch.equal(rd[q], wd[q]), q

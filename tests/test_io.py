"""Solution output / restart input (libmvd_io.so, include/mvd_io.h) on the host.

Pinned by:
  * the reference's own I/O test (io_test_main.cpp:114-227): every value codes its species
    and global (x, y, z) index as 0.sXXXyyyZZZ; the state is written as restart 5 at
    t = 8.5 and read back, and must agree to 1e-12 with the time to 1e-14. With unit
    factors 1 the round trip here is exact;
  * the file read without this library: h5dump (HDF5's own tool) lists the datasets, their
    types and shapes, and prints the values, which must be the code values times
    DensityUnits / MomentumUnits / EnergyUnits (euler3D.hpp:389-391) for the fluid fields
    and the species values unchanged, in [z][y][x] order (io.cpp:805-930);
  * restart_parameters.txt line by line against the reference's key order and formats
    (io.cpp:656-708, doubles as %.16e);
  * read_restart's compatibility checks (io.cpp:981-1050).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def io(mvd):
    if not os.path.exists(mvd.IO_LIB_PATH):
        pytest.fail("libmvd_io.so not built (run __graft_entry__.build())")
    mvd.io_lib()
    return mvd


def coded_state(nx, ny, nz, nchem):
    """io_test_main.cpp:124-151: value = 0.001 s + 1e-4 x + 1e-7 y + 1e-10 z."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    loc = 0.0001 * i + 0.0000001 * j + 0.0000000001 * k
    w = [np.ascontiguousarray((0.001 * v + loc).ravel()) for v in range(5)]
    if nchem:
        ch = np.stack([(0.001 * (5 + v) + loc).ravel() for v in range(nchem)], axis=1)
        w.append(np.ascontiguousarray(ch.ravel()))
    else:
        w.append(None)
    return w


def params(mvd, nx=7, ny=5, nz=4, nchem=10, units=(1.0, 1.0, 1.0)):
    p = mvd.make_params(nx, ny, nz, nchem, bc=(3, 3, 1, 1, 2, 2), gamma=5 / 3,
                        extent=((-1.0, 2.0), (0.0, 0.5), (0.25, 1.0)), units=units)
    p.cfl = 0.3
    return p


def h5dump_dataset(path, name):
    """(type, shape, flat values) of one dataset as h5dump prints it."""
    out = subprocess.run([H5DUMP, "-d", name, "-m", "%.17g", "-w", "0", "-y", path],
                         capture_output=True, text=True, check=True).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    sp = re.search(r"DATASPACE\s+(\w+)(?:\s*\{\s*\(([^)]*)\))?", out)
    shape = () if sp.group(1) == "SCALAR" else tuple(int(x) for x in sp.group(2).split(","))
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    vals = np.array([float(x) for x in re.split(r"[,\s]+", body.strip()) if x])
    return dtype, shape, vals


@pytest.mark.parametrize("nchem", [10, 0])
def test_reference_io_kat_round_trip(io, tmp_path, nchem):
    p = params(io, nchem=nchem)
    w = coded_state(7, 5, 4, nchem)
    io.output_solution(tmp_path, p, 8.5, 1e-3, 5, w, io.RunOptions())
    assert os.path.exists(tmp_path / "output-0000005.hdf5")
    assert os.path.exists(tmp_path / "restart_parameters.txt")
    r = [np.full_like(w[q], np.nan) for q in range(5)] + \
        ([np.full_like(w[5], np.nan)] if nchem else [None])
    t = io.read_restart(tmp_path, p, 5, r)
    assert abs(t - 8.5) <= 1e-14
    for q in range(6 if nchem else 5):
        assert np.array_equal(r[q], w[q]), q          # unit factors 1: exact


def test_file_contents_with_h5dump(io, tmp_path):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    units = (3.0e70, 3.0857e30, 1.0e11)
    p = params(io, units=units)
    w = coded_state(7, 5, 4, 10)
    io.output_solution(tmp_path, p, 8.5, 1e-3, 12, w)
    path = str(tmp_path / "output-0000012.hdf5")
    M, L, T = units
    U = [M / L / L / L] + [M / L / L / T] * 3 + [M / L / T / T]
    names = ["Density", "x-Momentum", "y-Momentum", "z-Momentum", "TotalEnergy"]
    for q, nm in enumerate(names):
        dtype, shape, vals = h5dump_dataset(path, "/" + nm)
        assert dtype == "H5T_IEEE_F64LE" and shape == (4, 5, 7), nm
        assert np.array_equal(vals, U[q] * w[q]), nm          # [z][y][x] = cell order
    ch = w[5].reshape(-1, 10)
    for v in range(10):
        dtype, shape, vals = h5dump_dataset(path, "/Chemical-%03d" % v)
        assert dtype == "H5T_IEEE_F64LE" and shape == (4, 5, 7)
        assert np.array_equal(vals, ch[:, v])
    dtype, shape, vals = h5dump_dataset(path, "/time")
    assert dtype == "H5T_IEEE_F64LE" and shape == () and vals.tolist() == [8.5]
    dtype, shape, vals = h5dump_dataset(path, "/nchem")
    assert dtype.startswith("H5T_STD_I32") and vals.tolist() == [10]
    dtype, shape, vals = h5dump_dataset(path, "/domain")
    assert shape == (3, 2) and vals.tolist() == [0.25, 1.0, 0.0, 0.5, -1.0, 2.0]
    # no parameter file without run options
    assert not os.path.exists(tmp_path / "restart_parameters.txt")
    # and back to code units: x * (1 / U) of the stored CGS value
    r = [np.empty_like(w[q]) for q in range(6)]
    assert io.read_restart(tmp_path, p, 12, r) == 8.5
    for q in range(5):
        assert np.array_equal(r[q], (1.0 / U[q]) * (U[q] * w[q])), q
        assert np.max(np.abs(r[q] - w[q]) / np.abs(w[q]).max()) <= 1e-12
    assert np.array_equal(r[5], w[5])


def test_restart_parameters_file(io, tmp_path):
    p = params(io, units=(3.0e70, 3.0857e30, 1.0e11))
    o = io.RunOptions(tf=10.0, nout=20, order=3, fixedstep=1, hmax=1e-4, etable=2)
    io.write_parameters(tmp_path, p, o, 1.25, 2.5e-5, 7)
    got = (tmp_path / "restart_parameters.txt").read_text().splitlines()
    e = "%.16e"
    want = ["# Euler3D restart file",
            "xl = " + e % -1.0, "xr = " + e % 2.0, "yl = " + e % 0.0, "yr = " + e % 0.5,
            "zl = " + e % 0.25, "zr = " + e % 1.0, "t0 = " + e % 1.25, "tf = " + e % 10.0,
            "gamma = " + e % (5 / 3), "MassUnits = " + e % 3.0e70,
            "LengthUnits = " + e % 3.0857e30, "TimeUnits = " + e % 1.0e11,
            "nx = 7", "ny = 5", "nz = 4", "xlbc = 3", "xrbc = 3", "ylbc = 1", "yrbc = 1",
            "zlbc = 2", "zrbc = 2", "cfl = " + e % 0.3, "nout = 13", "showstats = 0",
            "order = 3", "dense_order = -1", "etable = 2", "itable = -1", "mtable = -1",
            "adapt_method = 0", "maxnef = 0", "mxhnil = 0", "mxsteps = 5000",
            "safety = " + e % 0, "bias = " + e % 0, "growth = " + e % 0, "pq = 0",
            "k1 = " + e % 0, "k2 = " + e % 0, "k3 = " + e % 0, "etamx1 = " + e % 0,
            "etamxf = " + e % 0, "h0 = " + e % 2.5e-5, "hmin = " + e % 0,
            "hmax = " + e % 1e-4, "fixedstep = 1", "htrans = " + e % 0,
            "rtol = " + e % 1e-8, "atol = " + e % 1e-12, "fusedkernels = 1",
            "localreduce = 1", "restart = 7"]
    assert got == want


def test_read_restart_rejects_incompatible_files(io, tmp_path):
    p = params(io)
    w = coded_state(7, 5, 4, 10)
    io.output_solution(tmp_path, p, 8.5, 1e-3, 3, w)
    r = [np.empty_like(w[q]) for q in range(6)]
    with pytest.raises(io.MvdError) as e:                  # no such output
        io.read_restart(tmp_path, p, 4, r)
    assert e.value.code == io.MVD_IO_ERROR
    q = params(io, nchem=6)                                # species count differs
    r6 = [np.empty_like(w[k]) for k in range(5)] + [np.empty(7 * 5 * 4 * 6)]
    with pytest.raises(io.MvdError) as e:
        io.read_restart(tmp_path, q, 3, r6)
    assert e.value.code == io.MVD_IO_MISMATCH
    q = params(io)
    q.xr = 2.0 + 1e-12                                     # domain differs beyond 1e-14
    with pytest.raises(io.MvdError) as e:
        io.read_restart(tmp_path, q, 3, r)
    assert e.value.code == io.MVD_IO_MISMATCH
    q = params(io, nx=6)                                   # grid differs
    r_small = [np.empty(6 * 5 * 4) for _ in range(5)] + [np.empty(6 * 5 * 4 * 10)]
    with pytest.raises(io.MvdError) as e:
        io.read_restart(tmp_path, q, 3, r_small)
    assert e.value.code == io.MVD_IO_MISMATCH
    with pytest.raises(io.MvdError) as e:                  # species array missing
        io.output_solution(tmp_path, p, 0.0, 0.0, 1, w[:5] + [None])
    assert e.value.code == io.MVD_BAD_ARGUMENT


def test_io_library_exports_header_symbols(io):
    import ctypes
    L = ctypes.CDLL(io.IO_LIB_PATH)
    declared = io.exported_symbols_declared(io.IO_HEADER_PATH)
    assert declared == ["mvd_output_solution", "mvd_read_restart", "mvd_write_parameters"]
    assert all(hasattr(L, s) for s in declared)

"""Build libmvd_rhs.so (HIP kernels + C ABI) for gfx950, and libmvd_io.so (HDF5 solution
output / restart input, host code), in-tree.

    python sundials-manyvector-demo_amd/build.py [--jobs N]

Each .hip translation unit is compiled with hipcc for gfx950 only, then linked into
sundials-manyvector-demo_amd/libmvd_rhs.so. Objects go to sundials-manyvector-demo_amd/_build/.
csrc/io.cpp is compiled with g++ against the serial HDF5 of the image (HDF5_DIR, default
/opt/conda) and the HIP runtime into libmvd_io.so; without HDF5 it is skipped. arkode/
mvd_arkode.cpp (the ARKODE callbacks over the C ABI) is compiled with g++ into
libmvd_arkode.so.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmvd_rhs.so")
SOURCES = ["api.hip", "fe.hip", "fe_vjp.hip", "nvec.hip", "chem.hip", "chem_dirk.hip",
           "erk.hip", "mri.hip", "exchange.hip",
           "imex.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
IO_LIB = os.path.join(HERE, "libmvd_io.so")
ARKODE_LIB = os.path.join(HERE, "libmvd_arkode.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# -packed-fp32-ops: never form packed FP32 instructions (all arithmetic here is FP64).
FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-Wall",
         "-Wno-unused-function", "-Wno-unused-variable",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=fast-honor-pragmas", "-munsafe-fp-atomics"]


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err), src in zip(results, SOURCES):
            if err.strip():
                print(f"[{src}] {err.strip()}", file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_io()
    build_arkode()
    return LIB


def build_arkode():
    """libmvd_arkode.so: the ARKODE callbacks (include/mvd_arkode.h) over libmvd_rhs.so, host
    code. Built against the restated SUNDIALS declarations (arkode/mvd_sundials_min.h); its
    SUNDIALS symbols stay undefined and are resolved by the SUNDIALS libraries of the process
    that loads it (a driver linking SUNDIALS, or the tests' stand-in)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = ARKODE_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-I", inc, os.path.join(HERE, "arkode", "mvd_arkode.cpp"),
           "-o", tmp, "-L", HERE, "-l:libmvd_rhs.so", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libmvd_arkode.so build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ARKODE_LIB)
    return ARKODE_LIB


def build_io():
    """libmvd_io.so, or None when the HDF5 headers are absent."""
    if not os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")):
        print(f"libmvd_io.so skipped: no hdf5.h under {HDF5_DIR}/include", file=sys.stderr)
        return None
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = IO_LIB + ".tmp"
    # libstdc++ linked in statically: the HDF5 directory on the run path may hold an older
    # libstdc++ than the HIP runtime needs, and this library must not pull that one in
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-static-libstdc++", "-static-libgcc",
           "-I", inc, "-I", os.path.join(HDF5_DIR, "include"), "-I", os.path.join(ROCM, "include"),
           os.path.join(CSRC, "io.cpp"), "-o", tmp,
           "-L", os.path.join(HDF5_DIR, "lib"), "-lhdf5",
           "-Wl,-rpath," + os.path.join(HDF5_DIR, "lib"),
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libmvd_io.so build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))

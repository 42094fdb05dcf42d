"""Build libnmc.so (HIP kernels + C ABI) for gfx950 in-tree with hipcc.

Usage: python -m nanomagmc_amd.build   (or call build_lib()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnmc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["nmc_sweep.hip", "nmc_kernels.hip", "nmc_learn.hip", "nmc_api.hip"]
HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    return r


def build_lib(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


HDF5_PREFIX = os.environ.get("NMC_HDF5_PREFIX", "/opt/conda")
RUNNER = os.path.join(HERE, "nmc_run")


def build_runner(verbose=False):
    """The C++ host driver (nmc_run INPUT.dat) linked against libnmc.so and
    the serial HDF5 C library."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError("HDF5 headers not found under %s (set NMC_HDF5_PREFIX)" % HDF5_PREFIX)
    # The HDF5 prefix also holds an older libstdc++ than the ROCm runtime
    # needs, so neither -L nor an rpath may point there: libhdf5 is linked by
    # path and found at run time through an in-tree link directory.
    soname = "libhdf5.so.103"
    real = os.path.join(libdir, soname)
    if not os.path.exists(real):
        cands = sorted(f for f in os.listdir(libdir) if f.startswith("libhdf5.so."))
        if not cands:
            raise RuntimeError("no libhdf5 shared library in " + libdir)
        soname, real = cands[0], os.path.join(libdir, cands[0])
    linkdir = os.path.join(HERE, "hdf5")
    os.makedirs(linkdir, exist_ok=True)
    link = os.path.join(linkdir, soname)
    if os.path.lexists(link):
        os.remove(link)
    os.symlink(real, link)
    tmp = RUNNER + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(HERE, "host", "nmc_run.cpp"), "-o", tmp,
           "-I" + inc, "-L" + HERE, "-lnmc", real,
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/hdf5"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, RUNNER)
    return RUNNER


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_runner(verbose=True))

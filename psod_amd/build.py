"""Build recipe of the HIP library psod_amd/libpsod.so (gfx950) and of the
Fortran drop-in (psod_amd/fortran).  Compiles every time it is called."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsod.so")
FORTRAN = os.path.join(HERE, "fortran")

SOURCES = ["ctx.hip", "gen.hip", "cellsort.hip", "cic.hip", "host_api.hip",
           "tune.hip",
           "sod.hip", "sod_plain.hip", "greedy.hip", "pass.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("PSOD_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are disabled for every device pass
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def comp(s):
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, s), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> str | None:
    """Compile the ISO_C_BINDING module and the driver with flang (if the
    compiler exists) into psod_amd/fortran/psod_driver."""
    if not os.path.exists(FLANG) or not os.path.exists(
            os.path.join(FORTRAN, "psod_driver.f90")):
        return None
    out = os.path.join(FORTRAN, "psod_driver")
    mod_dir = os.path.join(BUILD, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    _run([FLANG, "-O2", "-module-dir", mod_dir,
          os.path.join(FORTRAN, "psod_sod.f90"),
          os.path.join(FORTRAN, "psod_driver.f90"),
          "-o", out, "-L" + HERE, "-lpsod", "-Wl,-rpath,$ORIGIN/.."])
    return out


HDF5 = os.environ.get("PSOD_HDF5", "/opt/conda")
H5LIB = os.path.join(HERE, "libpsod_h5.so")


def build_h5() -> str | None:
    """Compile the HDF5 halo writer (host C++, psod_amd/libpsod_h5.so) when
    the HDF5 C library is present."""
    inc = os.path.join(HDF5, "include", "hdf5.h")
    if not os.path.exists(inc):
        return None
    cxx = os.environ.get("CXX", "g++")
    tmp = H5LIB + ".tmp"
    _run([cxx, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra",
          "-I", os.path.join(HDF5, "include"),
          os.path.join(CSRC, "h5out.cpp"), "-o", tmp,
          "-L" + os.path.join(HDF5, "lib"), "-lhdf5",
          "-Wl,-rpath," + os.path.join(HDF5, "lib")])
    os.replace(tmp, H5LIB)
    return H5LIB


def build() -> str:
    lib = build_lib()
    build_fortran()
    build_h5()
    return lib


if __name__ == "__main__":
    print(build())
    shutil.rmtree(os.path.join(BUILD, "fmod"), ignore_errors=True)

"""Build the HIP library ``tpls_amd/libtpls_hip.so`` for gfx950 and the
Fortran host modules.

Every source under ``tpls_amd/csrc`` is compiled with hipcc in parallel and
linked into one shared library exporting the C ABI of
``include/tpls_core.h``.  ``tpls_amd/fortran/tpls_modules.f90`` (the modules
with the reference's names that main_ns_hybrid.F90 uses) is compiled with
amdflang into ``tpls_amd/fortran/libtpls_fortran.a`` plus its ``.mod`` files,
and the test driver ``tests/fortran/drive_modules`` is linked against both;
``tests/fortran/drive_io`` (Fortran list-directed I/O) is built for the
data-format tests.
Usage: ``python -m tpls_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtpls_hip.so")
FDIR = os.path.join(HERE, "fortran")
FMOD = os.path.join(FDIR, "mod")
FLIB = os.path.join(FDIR, "libtpls_fortran.a")
DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "drive_modules.f90")
DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_modules")
DRIVE_IO_SRC = os.path.join(ROOT, "tests", "fortran", "drive_io.f90")
DRIVE_IO = os.path.join(ROOT, "tests", "fortran", "drive_io")
DRIVE_IO_NATIVE = os.path.join(ROOT, "tests", "fortran", "drive_io_native")
DRIVE_MAIN = os.path.join(ROOT, "tests", "fortran", "drive_main_step")
FC = os.environ.get("TPLS_FC", "/opt/rocm/bin/amdflang")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    # packed FP32 instructions stay off: some of their forms have given wrong
    # results on these GPUs without any error; hipcc -O3 would emit them from
    # plain float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_fortran(verbose)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_fortran(verbose: bool = True) -> None:
    """The Fortran host modules and the test driver (skipped without amdflang)."""
    if not os.path.exists(FC):
        if verbose:
            print(f"{FC} not found: Fortran host modules not built")
        return
    os.makedirs(FMOD, exist_ok=True)
    obj = os.path.join(FMOD, "tpls_modules.o")
    _run([FC, "-O2", "-fPIC", "-module-dir", FMOD, "-c", os.path.join(FDIR, "tpls_modules.f90"), "-o", obj],
         "amdflang tpls_modules.f90")
    if os.path.exists(FLIB):
        os.remove(FLIB)
    _run(["ar", "rcs", FLIB, obj], "ar")
    dobj = os.path.join(FMOD, "drive_modules.o")
    _run([FC, "-O2", "-module-dir", FMOD, "-I", FMOD, "-c", DRIVER_SRC, "-o", dobj], "amdflang drive_modules.f90")
    _run([FC, "-o", DRIVER, dobj, FLIB, f"-L{HERE}", "-ltpls_hip", "-Wl,-rpath,$ORIGIN/../../tpls_amd"],
         "link drive_modules")
    # Fortran list-directed I/O in the data-file layout, and io_test.f90's
    # known answers through module tpls_io_native (tests/test_io.py)
    _run([FC, "-O2", "-module-dir", FMOD, "-o", DRIVE_IO, DRIVE_IO_SRC], "amdflang drive_io.f90")
    nobj = os.path.join(FMOD, "drive_io_native.o")
    _run([FC, "-O2", "-module-dir", FMOD, "-I", FMOD, "-c", DRIVE_IO_NATIVE + ".f90", "-o", nobj],
         "amdflang drive_io_native.f90")
    _run([FC, "-o", DRIVE_IO_NATIVE, nobj, FLIB, f"-L{HERE}", "-ltpls_hip", "-Wl,-rpath,$ORIGIN/../../tpls_amd"],
         "link drive_io_native")
    # main_ns_hybrid.F90's timestep walked call for call through the modules
    mobj = os.path.join(FMOD, "drive_main_step.o")
    _run([FC, "-O2", "-module-dir", FMOD, "-I", FMOD, "-c", DRIVE_MAIN + ".f90", "-o", mobj],
         "amdflang drive_main_step.f90")
    _run([FC, "-o", DRIVE_MAIN, mobj, FLIB, f"-L{HERE}", "-ltpls_hip", "-Wl,-rpath,$ORIGIN/../../tpls_amd"],
         "link drive_main_step")
    if verbose:
        print(f"built {FLIB} and {DRIVER}")


if __name__ == "__main__":
    build()
    sys.exit(0)

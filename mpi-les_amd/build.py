"""Build the gfx950 LES step library (lesamd/liblesamd.so) in-tree.

    python mpi-les_amd/build.py

Every HIP source is compiled for gfx950 with FMA contraction off (the FP32
stencils then round exactly like the reference's Fortran) and packed-FP32
instructions disabled, then linked into one shared library with a plain C
ABI (include/les_amd.h).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "lesamd" / "liblesamd.so"
SOURCES = ["k_step.hip", "k_press.hip", "k_output.hip", "k_noalloc.hip", "runtime.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LES_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", str(ROOT / "include"),
]


def _compile(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
# the reference's own Fortran flags that matter for the interface: free
# form, preprocessed, big-endian unformatted records (SConstruct:77)
FFLAGS = ["-cpp", "-O2", "-ffree-form", "-fPIC", "-fconvert=big-endian"]
FORTRAN_MODULES = HERE / "fortran" / "les_amd_modules.F90"
TIME_LOOP_SRC = ROOT / "tests" / "fortran" / "time_loop.F90"
TIME_LOOP_LIB = ROOT / "tests" / "fortran" / "liblestimeloop.so"


def build_fortran(verbose: bool = True) -> Path:
    """Compiles the drop-in Fortran modules (fortran/les_amd_modules.F90)
    and the test time loop against them into tests/fortran/liblestimeloop.so,
    linked to lesamd/liblesamd.so."""
    mod = OBJ / "fortran"
    mod.mkdir(parents=True, exist_ok=True)
    steps = [
        [FLANG, *FFLAGS, "-module-dir", str(mod), "-c", str(FORTRAN_MODULES), "-o", str(mod / "les_amd_modules.o")],
        [FLANG, *FFLAGS, "-module-dir", str(mod), "-c", str(TIME_LOOP_SRC), "-o", str(mod / "time_loop.o")],
        [FLANG, "-shared", "-fPIC", "-o", str(TIME_LOOP_LIB), str(mod / "les_amd_modules.o"), str(mod / "time_loop.o"),
         f"-L{LIB.parent}", "-llesamd", "-Wl,-rpath,$ORIGIN/../../mpi-les_amd/lesamd"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {TIME_LOOP_LIB.relative_to(ROOT)}")
    return TIME_LOOP_LIB


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)}")
    return LIB


if __name__ == "__main__":
    build()
    if Path(FLANG).exists():
        build_fortran()
    sys.exit(0)

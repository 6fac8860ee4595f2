"""Build the MI355X library ``libmeumapps_ss.so`` in-tree with hipcc (gfx950).

Every ``csrc/*.hip`` and ``csrc/*.cpp`` is compiled to an object under
``meumapps_ss_amd/_build`` (in parallel) and linked into the package
directory, so the built library travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmeumapps_ss.so")

ARCH = os.environ.get("MM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    drv = build_fortran_driver()
    if verbose and drv:
        print(f"built {drv}", file=sys.stderr)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(PKG, "fortran")
DRIVER = os.path.join(FDIR, "mm_step_driver")


def build_fortran_driver():
    """The Fortran step driver (fortran/mm_abi.f90 + fortran/mm_step_driver.f90)
    linked against the library with an rpath relative to itself; None where
    flang is absent."""
    if not os.path.exists(FLANG):
        return None
    mods = os.path.join(BUILD, "fmod")
    os.makedirs(mods, exist_ok=True)
    cmd = [FLANG, "-O2", "-J", mods, os.path.join(FDIR, "mm_abi.f90"), os.path.join(FDIR, "mm_step_driver.f90"),
           "-o", DRIVER + ".tmp", "-L", PKG, "-lmeumapps_ss", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=mods)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed for the step driver:\n{r.stderr[-6000:]}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build()

"""Build every native piece of the package in-tree (into ptb_amd/lib/).

* libptbgpu.so   - the HIP engine (ptb_amd/csrc/*.hip) for gfx950, C ABI of
                   include/ptbgpu.h;
* libptbsurr.so  - the (H, S) surrogate input generator (plain C, OpenMP);
* libptb_sm_gpu.so + sm_gpu_driver - the Fortran drop-in module `sm`
                   (ptb_amd/fortran/sm_gpu.f90) and a small Fortran driver that
                   calls it, built with flang when flang is available.

Run ``python -m ptb_amd.build``.  Every call compiles again.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
FORT = os.path.join(HERE, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_engine():
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    out = os.path.join(LIB, "libptbgpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_surrogate():
    out = os.path.join(LIB, "libptbsurr.so")
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "surrogate.c"), "-lm"])
    return out


def flang_path():
    for cand in ("/opt/rocm/bin/amdflang", shutil.which("flang") or ""):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran():
    """The Fortran drop-in (module sm over the C ABI) and its driver."""
    fc = flang_path()
    if fc is None or not os.path.exists(os.path.join(FORT, "sm_gpu.f90")):
        print("flang or fortran sources not available; skipping the Fortran shim")
        return None
    bdir = os.path.join(FORT, "_build")
    os.makedirs(bdir, exist_ok=True)
    flags = ["-fdefault-integer-8", "-O2", "-fPIC", "-module-dir", bdir]
    obj = os.path.join(bdir, "sm_gpu.o")
    _run([fc, *flags, "-c", os.path.join(FORT, "sm_gpu.f90"), "-o", obj])
    so = os.path.join(LIB, "libptb_sm_gpu.so")
    _run([fc, "-shared", "-o", so, obj, "-L", LIB, "-lptbgpu", "-Wl,-rpath,$ORIGIN"])
    drv = os.path.join(FORT, "sm_gpu_driver.f90")
    if os.path.exists(drv):
        exe = os.path.join(LIB, "sm_gpu_driver")
        _run([fc, *flags, "-o", exe, drv, "-L", LIB, "-lptb_sm_gpu", "-lptbgpu",
              "-Wl,-rpath,$ORIGIN"])
    return so


def build_all(fortran=True):
    os.makedirs(LIB, exist_ok=True)
    build_surrogate()
    build_engine()
    if fortran:
        build_fortran()


if __name__ == "__main__":
    build_all(fortran="--no-fortran" not in sys.argv)

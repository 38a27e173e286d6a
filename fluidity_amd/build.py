"""Build recipe of the gfx950 library (libflgpu.so) and of the Fortran shim.

Compiles every HIP source under fluidity_amd/csrc with hipcc for gfx950 and
links them into fluidity_amd/libflgpu.so (in-tree, so the built library
travels with the repository snapshot).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflgpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions (v_pk_*_f32) are disabled for all device code:
# some of their forms have given wrong results on this hardware.
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
    "-Wno-unused-command-line-argument",
]
# Per-file extra flags: the element arithmetic is written with explicit
# fused multiply-adds and must not be contracted any further.
EXTRA = {
    "assemble.hip": ["-ffp-contract=off"],
    "surface.hip": ["-ffp-contract=off"],
    "pressure.hip": ["-ffp-contract=off"],
}
SOURCES = ["flgpu_context.hip", "assemble.hip", "surface.hip", "krylov.hip", "learned.hip", "matrixdump.hip",
           "setup_gpu.hip", "pressure.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile and link libflgpu.so; returns its path."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose: bool = True) -> str | None:
    """Compile the ISO_C_BINDING module and the example driver with flang,
    when a Fortran compiler is present; returns the driver path or None."""
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    out = os.path.join(fdir, "flgpu_driver")
    os.makedirs(BUILD, exist_ok=True)
    cmd = [fc, "-O2", "-J", BUILD, "-o", out,
           os.path.join(fdir, "flgpu.F90"), os.path.join(fdir, "flgpu_driver.F90"),
           f"-L{HERE}", "-lflgpu", f"-Wl,-rpath,{HERE}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"flang failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
    if "--fortran" in sys.argv:
        build_fortran()

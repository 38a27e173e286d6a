"""petsc_readnsolve on the GPU: replay a matrix dump through the device
solver.

Fluidity writes a dump (DumpMatrixEquation, femtools/Petsc_Tools.F90:
1487-1504) when a solve fails to converge (femtools/Solvers.F90:1369-1479);
tools/petsc_readnsolve.F90 reads it back (:60-83, MatLoad/VecLoad) and
solves it again with the solver options of the command line.  This module
does the same with libflgpu.so: read the dump (flgpu_matrixdump_read), bind
its pattern, set its values and right-hand side, and solve on the device.

    python -m fluidity_amd.readnsolve matrixdump --ksp cg --pc jacobi --rtol 1e-8
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys

import numpy as np

from . import lib as L
from .cg import SolverOptions


def matrixdump_read(filename: str):
    """(findrm, colm, val, b, x0 or None), 1-based CSR."""
    lib = L.load()
    n, nnz = C.c_int32(0), C.c_int64(0)
    fn = filename.encode()
    L.check(lib.flgpu_matrixdump_read(fn, C.byref(n), C.byref(nnz), None, None, None, None, None, None),
            "flgpu_matrixdump_read")
    findrm = np.zeros(n.value + 1, np.int32)
    colm = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value)
    b = np.zeros(n.value)
    x0 = np.zeros(n.value)
    has = C.c_int32(0)
    L.check(lib.flgpu_matrixdump_read(fn, C.byref(n), C.byref(nnz), L.ptr(findrm), L.ptr(colm), L.ptr(val),
                                      L.ptr(b), L.ptr(x0), C.byref(has)), "flgpu_matrixdump_read")
    return findrm, colm, val, b, (x0 if has.value else None)


def matrixdump_write(filename: str, findrm, colm, val, b, x0=None) -> None:
    """DumpMatrixEquation's file for a 1-based CSR system."""
    lib = L.load()
    f = np.ascontiguousarray(findrm, np.int32)
    c = np.ascontiguousarray(colm, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    xx = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    L.check(lib.flgpu_matrixdump_write(filename.encode(), f.size - 1, L.ptr(f), L.ptr(c), L.ptr(v), L.ptr(bb),
                                       L.ptr(xx)), "flgpu_matrixdump_write")


def centrm_of(findrm: np.ndarray, colm: np.ndarray) -> np.ndarray:
    """Position (1-based) of each row's diagonal entry, 0 when absent."""
    n = findrm.size - 1
    row = np.repeat(np.arange(1, n + 1, dtype=np.int64), np.diff(findrm))
    cen = np.zeros(n, np.int32)
    hit = np.nonzero(colm == row)[0]
    cen[row[hit] - 1] = (hit + 1).astype(np.int32)
    return cen


class DumpSystem:
    """A dumped system bound to a device context."""

    def __init__(self, findrm, colm, val, b, device: int = 0):
        self.lib = L.load()
        self.ctx = C.c_void_p()
        L.check(self.lib.flgpu_init(C.byref(self.ctx), device), "flgpu_init")
        self.n = findrm.size - 1
        self.b = np.ascontiguousarray(b, np.float64)
        L.check(self.lib.flgpu_sparsity_bind(self.ctx, self.n, self.n, L.ptr(np.ascontiguousarray(findrm)),
                                             L.ptr(np.ascontiguousarray(colm)), L.ptr(centrm_of(findrm, colm))),
                "flgpu_sparsity_bind")
        L.check(self.lib.flgpu_matrix_set(self.ctx, L.ptr(np.ascontiguousarray(val, np.float64)), L.ptr(self.b),
                                          None), "flgpu_matrix_set")

    def solve(self, solver: SolverOptions, x0: np.ndarray | None = None):
        x = np.zeros(self.n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        info = L.SolveInfo()
        sp = solver.params()
        L.check(self.lib.flgpu_solve(self.ctx, C.byref(sp), L.ptr(x), C.byref(info)), "flgpu_solve")
        return x, info

    def close(self) -> None:
        if self.ctx:
            self.lib.flgpu_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def petsc_readnsolve(filename: str, solver: SolverOptions, zero_init_guess: bool = False, device: int = 0):
    """Read the dump and solve it on the device (the old-style path of
    tools/petsc_readnsolve.F90: the dumped initial guess unless
    zero_init_guess).  Returns (x, info)."""
    findrm, colm, val, b, x0 = matrixdump_read(filename)
    sysm = DumpSystem(findrm, colm, val, b, device)
    try:
        start_zero = zero_init_guess or x0 is None
        solver = SolverOptions(**{**solver.__dict__, "start_from_zero": start_zero})
        return sysm.solve(solver, None if start_zero else x0)
    finally:
        sysm.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("filename")
    ap.add_argument("--ksp", default="cg", choices=["cg", "gmres", "fcg"])
    ap.add_argument("--pc", default="jacobi", choices=["jacobi", "none", "learned"])
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--atol", type=float, default=0.0)
    ap.add_argument("--max-its", type=int, default=10000)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--zero-init-guess", action="store_true")
    a = ap.parse_args(argv)
    solver = SolverOptions(iterative_method=a.ksp, preconditioner=a.pc, relative_error=a.rtol,
                           absolute_error=a.atol, max_iterations=a.max_its, precision=a.precision)
    x, info = petsc_readnsolve(a.filename, solver, a.zero_init_guess)
    print(f"iterations: {info.its}  reason: {info.reason}  true relative residual: {info.true_relres:.3e}  "
          f"solve time: {info.seconds:.4f} s", flush=True)
    return 0 if info.reason > 0 else 1


if __name__ == "__main__":
    sys.exit(main())

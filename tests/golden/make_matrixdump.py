"""Writes tests/golden/c1_8.matrixdump, a dump in the layout
DumpMatrixEquation writes (femtools/Petsc_Tools.F90:1487-1504): the reduced
matrix (ghost pivots on the Dirichlet rows, Dirichlet columns dropped:
femtools/Petsc_Tools.F90:891-1074) that the oracle assembles for the 2-D
configuration on an 8 x 8 mesh (seed 1), the assembled rhs after the
Dirichlet step as right-hand side, and a zero initial guess.  Run from the
repository root: python tests/golden/make_matrixdump.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
import petsc_binary as PB  # noqa: E402
from fluidity_amd import meshgen as M  # noqa: E402


def main():
    cfg = M.config("c1", seed=1, scale=8)
    ref = O.reference_system(cfg)
    rowptr, col, val = O.reduced_system(ref["sparsity"], ref["A"], ref["inactive"])
    keep = val != 0.0
    row = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    lens = np.bincount(row[keep], minlength=rowptr.size - 1)
    findrm = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    colm = (col[keep] + 1).astype(np.int32)
    data = PB.encode(findrm, colm, val[keep], ref["rhs"], np.zeros(rowptr.size - 1))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_8.matrixdump")
    with open(out, "wb") as f:
        f.write(data)
    print(f"wrote {out}: n={rowptr.size - 1} nnz={int(keep.sum())} {len(data)} bytes")


if __name__ == "__main__":
    main()

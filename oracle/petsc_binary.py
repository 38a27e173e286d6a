"""TEST INFRASTRUCTURE ONLY: numpy restatement of PETSc's binary viewer
format for the matrix dumps Fluidity writes (DumpMatrixEquation,
femtools/Petsc_Tools.F90:1487-1504: MatView(A), VecView(b), VecView(x0)).

PETSc is not vendored by the reference (pinned 3.8.3 in debian/control:12,
3.12 in docker/Dockerfile.focal:30); its published binary layout with
32-bit indices is restated here: big-endian int32 MAT_FILE_CLASSID
1211216, rows, cols, nnz, row lengths, 0-based column indices, float64
values; each vector: int32 VEC_FILE_CLASSID 1211214, n, float64 values.
"""
from __future__ import annotations

import numpy as np

MAT_FILE_CLASSID = 1211216
VEC_FILE_CLASSID = 1211214


def encode(findrm, colm, val, b, x0=None) -> bytes:
    """Bytes of a dump of a 1-based CSR system."""
    findrm = np.asarray(findrm, np.int64)
    n = findrm.size - 1
    nnz = int(findrm[-1] - 1)
    parts = [np.array([MAT_FILE_CLASSID, n, n, nnz], ">i4").tobytes(),
             np.diff(findrm).astype(">i4").tobytes(),
             (np.asarray(colm, np.int64) - 1).astype(">i4").tobytes(),
             np.asarray(val, np.float64).astype(">f8").tobytes(),
             np.array([VEC_FILE_CLASSID, n], ">i4").tobytes(),
             np.asarray(b, np.float64).astype(">f8").tobytes()]
    if x0 is not None:
        parts += [np.array([VEC_FILE_CLASSID, n], ">i4").tobytes(),
                  np.asarray(x0, np.float64).astype(">f8").tobytes()]
    return b"".join(parts)


def decode(data: bytes):
    """(findrm, colm, val, b, x0 or None), 1-based CSR."""
    off = 0

    def take(dtype, count):
        nonlocal off
        size = np.dtype(dtype).itemsize * count
        if off + size > len(data):
            raise ValueError("truncated PETSc binary file")
        a = np.frombuffer(data, dtype=dtype, count=count, offset=off)
        off += size
        return a

    cid, m, ncols, nnz = (int(v) for v in take(">i4", 4))
    if cid != MAT_FILE_CLASSID:
        raise ValueError("not a PETSc binary matrix")
    lens = take(">i4", m).astype(np.int64)
    findrm = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    if findrm[-1] - 1 != nnz:
        raise ValueError("row lengths disagree with the header")
    colm = (take(">i4", nnz).astype(np.int64) + 1).astype(np.int32)
    val = take(">f8", nnz).astype(np.float64)
    cid, n = (int(v) for v in take(">i4", 2))
    if cid != VEC_FILE_CLASSID or n != m:
        raise ValueError("right-hand side missing")
    b = take(">f8", n).astype(np.float64)
    x0 = None
    if off < len(data):
        cid, n = (int(v) for v in take(">i4", 2))
        if cid != VEC_FILE_CLASSID or n != m:
            raise ValueError("bad initial guess")
        x0 = take(">f8", n).astype(np.float64)
    return findrm, colm, val, b, x0

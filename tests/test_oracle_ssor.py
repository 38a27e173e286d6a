"""Pins the oracle's multicolour SSOR (the "sor" preconditioner,
femtools/Solvers.F90:1844-1854 -> PCSOR, one symmetric Gauss-Seidel sweep,
omega = 1, taken colour by colour) with facts it does not compute itself:

* the row colouring against a brute-force first-fit colouring in row
  order, validity (no entry joins two rows of one colour) and the listing
  order;
* the sweep against the closed form M^-1 = (D + U)^-1 D (D + L)^-1 of the
  colour-permuted matrix (D, L, U its diagonal and strict triangles),
  formed with dense numpy solves;
* M^-1 symmetric for a symmetric matrix (the preconditioner CG needs);
* SSOR-PCG and SSOR-GMRES on the reduced systems of the configurations
  reach their tolerance, agree with a direct solve, and PCG needs fewer
  iterations than Jacobi-PCG.

PETSc's own SOR sweeps the rows in the natural order; its iteration path
is not reproduced (PETSc is not vendored), so iteration counts against
PETSc are parity unpinned.
"""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import oracle as O
from fluidity_amd import meshgen as M


def first_fit_rows(findrm, colm):
    n = findrm.size - 1
    colour = np.full(n, -1)
    for i in range(n):
        taken = {colour[j] for j in colm[findrm[i] - 1:findrm[i + 1] - 1] - 1 if j < i}
        q = 0
        while q in taken:
            q += 1
        colour[i] = q
    return colour


def reduced_full_pattern(ref):
    """The system petsc_solve hands to the Krylov method on the full pattern
    (ghost rows: the pivot on the diagonal; inactive columns zeroed), as the
    library and the oracle keep it: (findrm, colm, val, b)."""
    findrm, colm, centrm = ref["sparsity"]
    A, rhs, inact = ref["A"], ref["rhs"], ref["inactive"].astype(bool)
    n = findrm.size - 1
    row = np.repeat(np.arange(n), np.diff(findrm))
    col = colm - 1
    pivot = O.ghost_pivot(centrm, A)
    g = np.where(inact, rhs, 0.0)
    lift = np.zeros(n)
    np.add.at(lift, row, A * g[col])
    b = np.where(inact, 0.0, rhs - lift)
    val = np.where(inact[row], np.where(col == row, pivot, 0.0), np.where(inact[col], 0.0, A))
    return findrm, colm, val, b


@pytest.fixture(scope="module")
def c2_small():
    cfg = M.config("c2", seed=3, scale=4)
    return cfg, O.reference_system(cfg)


@pytest.mark.parametrize("name,scale", [("c1", 6), ("c2", 3), ("c3", 2)])
def test_row_colouring_is_first_fit_and_valid(name, scale):
    cfg = M.config(name, seed=2, scale=scale)
    findrm, colm, _ = O.reference_system(cfg)["sparsity"]
    ptr, rows = O.colour_rows(findrm, colm)
    brute = first_fit_rows(findrm, colm)
    n = findrm.size - 1
    colour = np.empty(n, int)
    for q in range(ptr.size - 1):
        members = rows[ptr[q]:ptr[q + 1]]
        assert np.all(np.diff(members) > 0)  # ascending inside a colour
        colour[members] = q
    assert np.array_equal(np.sort(rows), np.arange(n))
    assert np.array_equal(colour, brute)
    row = np.repeat(np.arange(n), np.diff(findrm))
    off = (colm - 1) != row
    assert not np.any(colour[row[off]] == colour[colm[off] - 1])


def dense_ssor_inverse(findrm, colm, val, r):
    n = findrm.size - 1
    A = sps.csr_matrix((val, colm - 1, findrm - 1), shape=(n, n)).toarray()
    ptr, rows = O.colour_rows(findrm, colm)
    perm = rows  # position k of the colour order holds row perm[k]
    B = A[np.ix_(perm, perm)]
    D = np.diag(np.diag(B))
    y = np.linalg.solve(np.tril(B), r[perm])   # (D + L) y = r
    zp = np.linalg.solve(np.triu(B), D @ y)    # (D + U) z = D y
    z = np.empty(n)
    z[perm] = zp
    return z


@pytest.mark.parametrize("seed", [0, 1])
def test_ssor_sweep_matches_closed_form(c2_small, seed):
    cfg, ref = c2_small
    findrm, colm, val, _ = reduced_full_pattern(ref)
    r = np.random.default_rng(seed).standard_normal(findrm.size - 1)
    z = O.ssor_apply(findrm, colm, val, r)
    zd = dense_ssor_inverse(findrm, colm, val, r)
    assert np.max(np.abs(z - zd)) <= 1e-12 * np.max(np.abs(zd))


def test_ssor_sweep_nonsymmetric_closed_form():
    """The advection-diffusion matrix of C1 (non-symmetric): the same closed
    form (the sweep reads rows, so a transposed operand would fail it)."""
    cfg = M.config("c1", seed=4, scale=6)
    ref = O.reference_system(cfg)
    findrm, colm, val, _ = reduced_full_pattern(ref)
    n = findrm.size - 1
    assert not np.allclose(sps.csr_matrix((val, colm - 1, findrm - 1), shape=(n, n)).toarray(),
                           sps.csr_matrix((val, colm - 1, findrm - 1), shape=(n, n)).toarray().T)
    r = np.random.default_rng(9).standard_normal(n)
    z = O.ssor_apply(findrm, colm, val, r)
    zd = dense_ssor_inverse(findrm, colm, val, r)
    assert np.max(np.abs(z - zd)) <= 1e-12 * np.max(np.abs(zd))


def test_ssor_inverse_is_symmetric(c2_small):
    cfg, ref = c2_small
    findrm, colm, val, _ = reduced_full_pattern(ref)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, findrm.size - 1))
    a = u @ O.ssor_apply(findrm, colm, val, v)
    b = v @ O.ssor_apply(findrm, colm, val, u)
    assert abs(a - b) <= 1e-12 * max(abs(a), 1e-300)


@pytest.mark.parametrize("name,scale,ksp", [("c2", 5, "cg"), ("c1", 10, "gmres"), ("c3", 3, "gmres")])
def test_ssor_krylov_matches_direct_solve(name, scale, ksp):
    cfg = M.config(name, seed=6, scale=scale)
    ref = O.reference_system(cfg)
    findrm, colm, val, b = reduced_full_pattern(ref)
    n = findrm.size - 1
    x, its, reason, relres, _ = O.petsc_solve(ksp, "sor", 1e-10, ref["sparsity"], ref["A"], ref["inactive"],
                                              ref["rhs"], max_its=5000)
    assert reason > 0, reason
    A = sps.csr_matrix((val, colm - 1, findrm - 1), shape=(n, n)).tocsc()
    y = spla.spsolve(A, b)
    xd = np.where(ref["inactive"].astype(bool), ref["rhs"], y)
    assert np.max(np.abs(x - xd)) <= 1e-7 * np.max(np.abs(xd))
    _, its_j, reason_j, _, _ = O.petsc_solve(ksp, "jacobi", 1e-10, ref["sparsity"], ref["A"], ref["inactive"],
                                             ref["rhs"], max_its=5000)
    assert reason_j > 0
    if ksp == "cg":
        assert its < its_j, (its, its_j)

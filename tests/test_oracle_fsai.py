"""The oracle's FSAI factor (the ceiling of the learned preconditioner's
factor family) pinned by what the mathematics fixes:

* on the full lower pattern of a dense SPD matrix the factor is exact:
  G^T G = A^-1 (G is the inverse Cholesky factor of A, with the rows
  reversed in order of elimination);
* on a diagonal pattern it is Jacobi, G = diag(A)^-1/2;
* on a sparse pattern its defining conditions hold: (G A G^T)_ii = 1 and
  (G A)_ij = 0 for j in the row's pattern, j != i (checked with a dense
  product, not the local solves the factor is built from);
* its Kaporin number is no larger than Jacobi's (FSAI minimises it over
  the pattern), and a row over kmax keeps the Jacobi row.
"""
import numpy as np

import oracle as O


def csr(A):
    n = A.shape[0]
    rowptr = [0]
    col, val = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        col.extend(nz)
        val.extend(A[i, nz])
        rowptr.append(len(col))
    return np.array(rowptr), np.array(col), np.array(val)


def dense_G(G, n):
    out = np.zeros((n, n))
    for i, (c, v) in G.items():
        out[i, c] = v
    return out


def spd(n, seed, density=1.0):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) * (rng.random((n, n)) < density)
    A = B @ B.T
    A[np.abs(A) < 1e-12] = 0.0
    return A + n * np.eye(n)


def laplacian_2d(m, contrast=1.0, seed=0):
    n = m * m
    rng = np.random.default_rng(seed)
    k = np.exp(np.log(contrast) * rng.random(n))
    A = np.zeros((n, n))
    for i in range(m):
        for j in range(m):
            p = i * m + j
            for di, dj in ((1, 0), (0, 1)):
                ii, jj = i + di, j + dj
                if ii < m and jj < m:
                    q = ii * m + jj
                    w = 0.5 * (k[p] + k[q])
                    A[p, p] += w
                    A[q, q] += w
                    A[p, q] -= w
                    A[q, p] -= w
    return A + 1e-3 * np.eye(n)


def kaporin(M):
    """(trace / n) / det^(1/n) of an SPD matrix."""
    n = M.shape[0]
    sign, logdet = np.linalg.slogdet(M)
    assert sign > 0
    return np.trace(M) / n / np.exp(logdet / n)


def test_full_pattern_is_the_exact_inverse():
    A = spd(12, 1)
    G = dense_G(O.fsai_factor(*csr(A)), 12)
    assert np.allclose(G, np.tril(G))
    assert np.max(np.abs(G.T @ G - np.linalg.inv(A))) <= 1e-12 * np.max(np.abs(np.linalg.inv(A)))
    # the same factor as the inverse of the Cholesky factor: A = L L^T gives
    # A^-1 = L^-T L^-1 with L^-1 lower triangular
    Gc = np.linalg.inv(np.linalg.cholesky(A))
    assert np.max(np.abs(G - Gc)) <= 1e-12 * np.max(np.abs(Gc))


def test_diagonal_pattern_is_jacobi():
    A = np.diag(np.linspace(1.0, 5.0, 7))
    G = dense_G(O.fsai_factor(*csr(A)), 7)
    assert np.allclose(np.diag(G), 1.0 / np.sqrt(np.diag(A)), rtol=1e-15, atol=0)
    assert np.count_nonzero(G - np.diag(np.diag(G))) == 0


def test_sparse_pattern_conditions_and_kaporin():
    A = laplacian_2d(7, contrast=1e3, seed=2)
    n = A.shape[0]
    rowptr, col, val = csr(A)
    Gd = O.fsai_factor(rowptr, col, val)
    G = dense_G(Gd, n)
    GAG = G @ A @ G.T
    assert np.max(np.abs(np.diag(GAG) - 1.0)) <= 1e-12
    GA = G @ A
    for i, (c, v) in Gd.items():
        off = c[c != i]
        assert np.max(np.abs(GA[i, off]), initial=0.0) <= 1e-12 * np.max(np.abs(A))
        assert set(c) == set(col[rowptr[i]:rowptr[i + 1]][col[rowptr[i]:rowptr[i + 1]] <= i])
    Dj = np.diag(1.0 / np.sqrt(np.diag(A)))
    assert kaporin(GAG) < kaporin(Dj @ A @ Dj)
    # apply: z = G^T G r
    r = np.random.default_rng(3).standard_normal(n)
    assert np.max(np.abs(O.fsai_apply(Gd, r) - G.T @ (G @ r))) <= 1e-13 * np.max(np.abs(G.T @ (G @ r)))


def test_rows_over_kmax_keep_jacobi():
    A = spd(10, 4)
    Gd = O.fsai_factor(*csr(A), kmax=4)
    for i in range(10):
        c, v = Gd[i]
        if i + 1 > 4:
            assert list(c) == [i] and v[0] == 1.0 / np.sqrt(A[i, i])
        else:
            assert c.size == i + 1


def test_nonsymmetric_matrix_reads_the_lower_triangle():
    A = laplacian_2d(5, seed=6)
    N = A.copy()
    N[np.triu_indices(25, 1)] *= 1.5  # upper triangle changed, lower kept
    N[np.abs(A) == 0] = 0.0
    G1, G2 = O.fsai_factor(*csr(A)), O.fsai_factor(*csr(N))
    for i in range(25):
        assert np.array_equal(G1[i][0], G2[i][0]) and np.array_equal(G1[i][1], G2[i][1])

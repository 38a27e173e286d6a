"""The oracle's pressure path (divergence block ct_m, lumped velocity mass,
C^T M_l^-1 C) pinned by what the mathematics fixes, on P1-P1 and
Taylor-Hood P2-P1 meshes in 2-D and 3-D:

* P1 velocity, P1 pressure: each element block is the closed form
  |T|/(d+1) grad N_j (the integral of a P1 pressure function is |T|/(d+1),
  the velocity gradient is constant);
* the discrete divergence of a linear velocity u = B x + c (in both velocity
  spaces) is tr(B) times the integral of each pressure basis function
  (divergence constant, the quadrature exact for it);
* by parts and not by parts agree on every row of an interior pressure node
  (the boundary term vanishes where N_i does) and differ on boundary rows;
* the lumped masses sum to the integral of the density (partition of unity);
* C^T M_l^-1 C equals a dense product computed another way, and is
  symmetric bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import meshgen as M

MESHES = [("p1p1-2d", lambda: M.square_tris(6, jitter=0.2, seed=3)),
          ("p1p1-3d", lambda: M.cube_tets(4, jitter=0.15, seed=4)),
          ("p2p1-2d", lambda: M.square_tris(5, jitter=0.2, seed=5, degree=2)),
          ("p2p1-3d", lambda: M.cube_tets(3, jitter=0.15, seed=6, degree=2))]


def quad_degree(mesh):
    return 2 if mesh.degree == 1 else 4


def dense_ct(pd, ct_sp, ct_val, p_nodes, u_nodes):
    f, c = ct_sp
    dim = ct_val.shape[0]
    out = np.zeros((dim, p_nodes, u_nodes))
    rows = np.repeat(np.arange(p_nodes), np.diff(f))
    for d in range(dim):
        out[d, rows, c - 1] = ct_val[d]
    return out


def vertex_positions(mesh):
    return np.asarray(mesh.X, np.float64).reshape(mesh.nx, mesh.dim)


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_divergence_of_linear_field(name, make):
    mesh = make()
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, quad_degree(mesh))
    ct, ml = O.ct_assemble(asm, pd)
    d = mesh.dim
    rng = np.random.default_rng(1)
    B = rng.standard_normal((d, d))
    c0 = rng.standard_normal(d)
    u = mesh.node_pos @ B.T + c0  # (nodes, dim): exact in P1 and P2
    D = dense_ct(pd, cts, ct, p_nodes, mesh.nnodes)
    div = sum(D[a] @ u[:, a] for a in range(d))
    # integral of each vertex basis function: |T|/(d+1) per element
    xv = vertex_positions(mesh)
    ex = np.asarray(mesh.ndglno_x, np.int64).reshape(-1, d + 1) - 1
    vol = np.abs(np.linalg.det(xv[ex[:, :d]] - xv[ex[:, d:d + 1]])) / (2 if d == 2 else 6)
    integ = np.zeros(p_nodes)
    np.add.at(integ, ex.ravel(), np.repeat(vol / (d + 1), d + 1))
    expect = np.trace(B) * integ
    assert np.max(np.abs(div - expect)) <= 1e-12 * np.max(np.abs(expect))


@pytest.mark.parametrize("name,make", MESHES[:2], ids=[m[0] for m in MESHES[:2]])
def test_p1p1_element_closed_form(name, make):
    mesh = make()
    asm, pd, *_ = O.pressure_setup(mesh, quad_degree(mesh))
    d = mesh.dim
    xv = vertex_positions(mesh)
    ex = np.asarray(mesh.ndglno_x, np.int64).reshape(-1, d + 1) - 1
    for ele in (0, mesh.nele // 2, mesh.nele - 1):
        mat, _ = O.ct_element(asm, pd, ele)
        x = xv[ex[ele]]
        J = (x[:d] - x[d]).T  # columns: edge vectors to the last vertex
        vol = abs(np.linalg.det(J)) / (2 if d == 2 else 6)
        Ginv = np.linalg.inv(J)  # barycentric gradients: rows of J^-1 for the first d vertices
        G = np.vstack([Ginv, -Ginv.sum(0)])  # (d+1, d)
        expect = np.einsum("ja->aj", G)[:, None, :] * np.ones((1, d + 1, 1)) * (vol / (d + 1))
        assert np.max(np.abs(mat - expect)) <= 1e-13 * np.max(np.abs(expect))


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_by_parts_agrees_on_interior_rows(name, make):
    mesh = make()
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, quad_degree(mesh), ibp=False)
    ct, _ = O.ct_assemble(asm, pd)
    asm2, pd2, *_ = O.pressure_setup(mesh, quad_degree(mesh), ibp=True)
    ct2, _ = O.ct_assemble(asm2, pd2)
    f, c = cts
    rows = np.repeat(np.arange(p_nodes), np.diff(f))
    xv = vertex_positions(mesh)
    boundary = np.any((xv < 1e-12) | (xv > 1 - 1e-12), axis=1)
    inner = ~boundary[rows]
    scale = np.max(np.abs(ct))
    assert np.max(np.abs(ct[:, inner] - ct2[:, inner])) <= 1e-13 * scale
    assert np.max(np.abs(ct[:, ~inner] - ct2[:, ~inner])) > 1e-6 * scale


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_lumped_mass_sums_to_the_density_integral(name, make):
    mesh = make()
    _, _, *_ = O.pressure_setup(mesh, quad_degree(mesh))
    rho = 1.0 + mesh.node_pos[:, 0]  # linear: exact in both spaces
    asm, pd, *_ = O.pressure_setup(mesh, quad_degree(mesh), density=rho)
    _, ml = O.ct_assemble(asm, pd)
    assert abs(ml.sum() - 1.5) <= 1e-13  # integral of 1 + x over the unit square / cube
    asm1, pd1, *_ = O.pressure_setup(mesh, quad_degree(mesh), density=2.0)
    _, ml1 = O.ct_assemble(asm1, pd1)
    assert abs(ml1.sum() - 2.0) <= 1e-13
    assert np.all(ml1 > 0) or mesh.degree == 2  # P2 row sums of a consistent mass can be negative


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_cmc_against_dense_product(name, make):
    mesh = make()
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, quad_degree(mesh))
    ct, ml = O.ct_assemble(asm, pd)
    d = mesh.dim
    rng = np.random.default_rng(2)
    inv = rng.uniform(0.5, 2.0, size=(d, mesh.nnodes))
    cs = O.cmc_sparsity(*cts, p_nodes, mesh.nnodes)
    val = O.cmc(cts, ct, inv, cs)
    D = dense_ct(pd, cts, ct, p_nodes, mesh.nnodes)
    dense = sum(D[a] @ np.diag(inv[a]) @ D[a].T for a in range(d))
    f, c = cs
    rows = np.repeat(np.arange(p_nodes), np.diff(f))
    assert np.max(np.abs(val - dense[rows, c - 1])) <= 1e-13 * np.max(np.abs(dense))
    # nothing of the product falls outside the pattern
    mask = np.zeros_like(dense, bool)
    mask[rows, c - 1] = True
    assert np.max(np.abs(dense[~mask]), initial=0.0) == 0.0
    # symmetric bit for bit (a b c = b a c)
    full = np.zeros_like(dense)
    full[rows, c - 1] = val
    assert np.array_equal(full, full.T)


def test_out_of_pattern_is_an_error():
    mesh = M.cube_tets(2, jitter=0.1, seed=7)
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, 2)
    f, c = cts
    keep = np.ones(c.size, bool)
    keep[f[1] - 1 - 1] = False  # the last entry of row 1
    f2 = f.copy()
    f2[1:] -= 1
    bad = (f2, c[keep].copy())
    from fluidity_amd.elements import make_element
    pd2 = O.PressureData(pn, mesh.dim + 1, p_nodes, make_element(mesh.dim, 1, 2), bad, 1.0, False)
    with pytest.raises(RuntimeError, match="outside"):
        O.ct_assemble(asm, pd2)

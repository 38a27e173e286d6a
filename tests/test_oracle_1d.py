"""Pins the oracle's 1-D (interval) path.

* The reference's own known-answer case, tests/diff_src_dirichlet_1d
  (diff_src_dirichlet_1d_cg_p1_A.flml): -T'' = 2.13 on [0, 10] with dx = 1,
  T(10) = 2, zero flux at x = 0, no mass or advection terms, theta = 1,
  dt = 5, GMRES + sor to 1e-10.  Its pass test asks for the nodal error
  against T = 2.13/2 (100 - x^2) + 2 to stay below 1e-8 (P1 in 1-D is
  nodally exact for this problem on any mesh, P2 exact everywhere: the
  solution is quadratic), here also on a jittered mesh and in P2.
* Interval quadrature integrates monomials exactly; closed-form P1
  interval element matrices: mass h/6 [2 1; 1 2], stiffness kbar/h
  [1 -1; -1 1], advection u/2 [-1 1; -1 1] (constant u, beta = 0).
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import elements as E
from fluidity_amd import meshgen as M


def dsd1d_solve(cfg):
    """The oracle's step of the known-answer case: assembly, Dirichlet value
    2 at the right node, GMRES + sor, T1 = T0 + dt dT."""
    ref = O.reference_system(cfg)
    rhs = ref["rhs"].copy()
    bn = cfg.mesh.boundary - 1
    rhs[bn] = (2.0 - cfg.T[bn]) / cfg.dt      # the value form of the Dirichlet row
    x, its, reason, _, _ = O.petsc_solve("gmres", "sor", 1e-10, ref["sparsity"], ref["A"], ref["inactive"], rhs,
                                         atol=1e-10)
    assert reason > 0
    return cfg.T + cfg.dt * x


def dsd1d_exact(x):
    return 2.13 / 2.0 * (100.0 - x ** 2) + 2.0


@pytest.mark.parametrize("name,n,jitter", [("dsd1d", 10, 0.0), ("dsd1d", 37, 0.3), ("dsd1dp2", 10, 0.0),
                                           ("dsd1dp2", 23, 0.3)])
def test_diff_src_dirichlet_1d_known_answer(name, n, jitter):
    cfg = M.config(name, scale=n)
    if jitter:
        deg = cfg.mesh.degree
        cfg.mesh = M.interval_mesh(n, 0.0, 10.0, jitter=jitter, seed=5, degree=deg)
        cfg.mesh.boundary = np.array([cfg.mesh.nnodes if deg == 1 else 2 * n + 1], np.int32)
        cfg.T = np.zeros(cfg.mesh.nnodes)
        cfg.source = np.full(cfg.mesh.nnodes, 2.13)
    T1 = dsd1d_solve(cfg)
    err = np.abs(T1 - dsd1d_exact(cfg.mesh.node_pos[:, 0]))
    assert err.max() < 1e-8, err.max()


def test_known_answer_fails_a_planted_sign_error():
    """The same step with the source's sign flipped misses the answer by
    far more than the tolerance (the check can fail)."""
    cfg = M.config("dsd1d")
    cfg.source = -cfg.source
    T1 = dsd1d_solve(cfg)
    assert np.abs(T1 - dsd1d_exact(cfg.mesh.node_pos[:, 0])).max() > 1.0


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_interval_quadrature_integrates_monomials_exactly(degree):
    """int_0^1 x^a dx = 1/(a+1) for a <= the rule's degree (the rules'
    literal constants carry about 16 digits, so to 1e-14)."""
    l, w = O.quadrature(1, degree)
    for a in range(degree + 1):
        assert abs(np.dot(w, l[:, 0] ** a) - 1.0 / (a + 1)) < 1e-14, (degree, a)


def test_interval_tables_agree_with_package_and_interpolate():
    for deg, qd in [(1, 2), (2, 3), (2, 4)]:
        l, w = O.quadrature(1, qd)
        l2, w2 = E.quadrature(1, qd)
        n, dn = O.shape_tables(1, deg, l)
        n2, dn2 = E.shape_tables(1, deg, l2)
        assert np.array_equal(l, l2) and np.array_equal(w, w2)
        assert np.array_equal(n, n2) and np.array_equal(dn, dn2)
    counts = E.simplex_counts(1, 2)
    n, _ = O.shape_tables(1, 2, np.ascontiguousarray(counts / 2.0))
    assert np.allclose(n, np.eye(3), atol=1e-15)


def _interval_assembly(m, T, U, kap, src, p):
    el = E.make_element(1, m.degree, 2)
    xel = E.make_element(1, 1, 2)
    fields = {"T": T, "U": U, "kappa": kap, "src": src, "absorp": None}
    sp = O.make_sparsity(m)
    col, ncol = O.colour_elements(m)
    return O.Assembly(m, el, xel, fields, sp, O.colour_sets(col, ncol))


def test_p1_interval_element_closed_forms():
    m = M.interval_mesh(9, -1.0, 2.0, jitter=0.4, seed=2)
    N = m.nnodes
    kap = 0.5 + 0.25 * m.node_pos[:, 0] ** 2
    u = 0.7
    for what in ("mass", "stiffness", "advection"):
        p = O.params(1.0, 1.0, 0.0, have_mass=what == "mass", adv=what == "advection",
                     diff=what == "stiffness", src=False)
        K = np.zeros((N, 1, 1))
        K[:, 0, 0] = kap
        asm = _interval_assembly(m, np.zeros(N), np.full((N, 1), u), K if what == "stiffness" else None, None, p)
        for e in [0, 4, 8]:
            mat, _ = O.element(p, asm, e)
            a, b = m.ndglno[2 * e] - 1, m.ndglno[2 * e + 1] - 1
            h = m.X[b, 0] - m.X[a, 0]
            if what == "mass":
                want = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
            elif what == "stiffness":
                want = 0.5 * (kap[a] + kap[b]) / h * np.array([[1.0, -1.0], [-1.0, 1.0]])
            else:
                want = u / 2.0 * np.array([[-1.0, 1.0], [-1.0, 1.0]])
            assert np.allclose(mat, want, rtol=1e-13, atol=1e-15), (what, e, mat, want)


def _interval_surface_setup(m, T, U, kappa, ibp, beta=1.0, qd=2):
    from fluidity_amd.meshgen import surface_of, unit_box_bcs
    el = E.make_element(1, m.degree, qd)
    xel = E.make_element(1, 1, qd)
    K = None if kappa is None else np.full((m.nnodes, 1, 1), kappa)
    fields = {"T": T, "U": U, "kappa": K, "src": None, "absorp": None}
    sp = O.make_sparsity(m)
    col, ncol = O.colour_elements(m)
    asm = O.Assembly(m, el, xel, fields, sp, O.colour_sets(col, ncol))
    surf = surface_of(m)
    bt, b1, b2, strong = unit_box_bcs(m, surf, "neumann_robin")
    sd = O.SurfaceData(m, surf, E.make_face_element(1, m.degree, qd), E.make_face_element(1, 1, qd), bt, b1, b2)
    p = O.params(0.1, 0.5, beta, have_mass=True, adv=U is not None, ibp=ibp, diff=kappa is not None)
    return asm, sd, surf, p, (bt, b1, b2)


@pytest.mark.parametrize("degree", [1, 2])
def test_interval_point_faces_neumann_robin(degree):
    """A point facet has detwei 1 and N = 1: Neumann adds g to the end
    node's rhs; Robin adds dt theta h to its diagonal and g - h T to its
    rhs (Advection_Diffusion_CG.F90:1349-1379)."""
    m = M.interval_mesh(7, 0.0, 1.0, jitter=0.2, seed=1, degree=degree)
    T = np.cos(2 * m.node_pos[:, 0])
    asm, sd, surf, p, (bt, b1, b2) = _interval_surface_setup(m, T, None, 0.7, ibp=False)
    A0, r0, _ = O.assemble(p, asm)
    A1, r1 = A0.copy(), r0.copy()
    O.surface_assemble(p, asm, sd, A1, r1)
    f, c, _ = O.make_sparsity(m)
    n0, n1 = 0, m.nnodes - 1 if degree == 1 else 2 * 7
    left = np.nonzero(bt == M.BC_NEUMANN)[0]  # x = 0
    assert left.size == 1 and surf.face_nodes[left[0]] == n0 + 1
    dr = r1 - r0
    g0 = b1[left[0]]
    assert abs(dr[n0] - g0) < 1e-15
    right = [i for i in range(surf.nfaces) if i != left[0]][0]
    g1, h1 = b1[right], b2[right]
    assert abs(dr[n1] - (g1 - h1 * T[n1])) < 1e-14
    diag = [k for k in range(f[n1] - 1, f[n1 + 1] - 1) if c[k] == n1 + 1][0]
    assert abs((A1 - A0)[diag] - 0.1 * 0.5 * h1) < 1e-15
    assert np.count_nonzero(A1 - A0) == 1 and np.count_nonzero(dr) == 2


@pytest.mark.parametrize("degree", [1, 2])
def test_interval_integration_by_parts_identity(degree):
    """Non-by-parts (beta = 1) = by parts + the point faces' u.n T term
    (rules exact for the integrands: degree 2 for P1, 4 for P2)."""
    m = M.interval_mesh(9, 0.0, 1.0, jitter=0.3, seed=2, degree=degree)
    x = m.node_pos[:, 0]
    T = np.sin(3 * x) + x
    U = (0.5 + 0.3 * x)[:, None]
    out = {}
    for ibp in (False, True):
        asm, sd, surf, p, _ = _interval_surface_setup(m, T, U, None, ibp=ibp, qd=2 * degree)
        sd.s.bc_type = O._p(np.zeros(surf.nfaces, np.int32))
        A, r, _ = O.assemble(p, asm)
        if ibp:
            O.surface_assemble(p, asm, sd, A, r)
        out[ibp] = (A, r)
    sa = np.abs(out[False][0]).max()
    assert np.abs(out[True][0] - out[False][0]).max() < 1e-13 * sa
    assert np.abs(out[True][1] - out[False][1]).max() < 1e-13 * np.abs(out[False][1]).max()

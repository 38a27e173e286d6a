"""Pins the oracle's mesh-movement branch of the element routine
(assemble/Advection_Diffusion_CG.F90:421-437 options, :788-791 detwei of
the old and new positions, :884-938 the mass at the new positions and the
(N^{n+1} - N^n) T^n / dt rhs term, :983-985 the velocity relative to the
grid, :1099-1102 the grid velocity's divergence term).

* A mesh that does not move (old = new = current positions, zero grid
  velocity) gives the static assembly.
* A uniform dilation by (1+s): the mass is (1+s)^d times the static one
  and the rhs gains -((1+s)^d - 1) M T / dt (consistent and lumped).
* A rigid translation with the matching grid velocity: the static assembly
  with the velocity relative to the grid (non-by-parts and by-parts).
* A linear grid velocity equal to the flow velocity: the advection matrix
  is -div(u_grid) M (non-by-parts; the by-parts form has no such term).
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import elements as E
from fluidity_amd import meshgen as M


def setup(dim, seed=4):
    m = M.square_tris(5, jitter=0.25, seed=seed) if dim == 2 else M.cube_tets(3, jitter=0.2, seed=seed)
    p = m.node_pos
    T = np.cos(2 * p[:, 0]) + p[:, 1] ** 2
    U = np.stack([np.sin(p[:, 1] + d) for d in range(dim)], 1) * 0.7
    return m, T, U


def assemble(m, T, U, *, kappa=0.3, ibp=False, beta=0.0, lump=False, adv=True, X_old=None, X_new=None, ug=None,
             move=False, dt=0.1):
    el = E.make_element(m.dim, 1, 2)
    kap = (np.eye(m.dim) * kappa).ravel() if kappa is not None else None
    fields = {"T": T, "U": U, "kappa": kap, "src": None, "absorp": None,
              "X_old": X_old, "X_new": X_new, "ug": ug}
    sp = O.make_sparsity(m)
    col, ncol = O.colour_elements(m)
    asm = O.Assembly(m, el, el, fields, sp, O.colour_sets(col, ncol))
    p = O.params(dt, 0.5, beta, have_mass=True, lump=lump, adv=adv, ibp=ibp, diff=kappa is not None,
                 move_mesh=move)
    A, rhs, _ = O.assemble(p, asm)
    return A, rhs, sp


def csr_matvec(sp, A, x):
    f, c, _ = sp
    rows = np.repeat(np.arange(f.size - 1), np.diff(f))
    return np.bincount(rows, A * x[c - 1], minlength=f.size - 1)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("dim", [2, 3])
def test_still_mesh_is_static_assembly(dim):
    m, T, U = setup(dim)
    X = m.X.copy()
    for ibp in (False, True):
        A0, r0, _ = assemble(m, T, U, ibp=ibp, beta=0.4)
        A1, r1, _ = assemble(m, T, U, ibp=ibp, beta=0.4, X_old=X, X_new=X, ug=np.zeros_like(X), move=True)
        assert np.array_equal(A0, A1) and np.array_equal(r0, r1)


@pytest.mark.parametrize("dim,lump", [(2, False), (3, False), (3, True)])
def test_dilation_scales_mass_and_adds_rhs_term(dim, lump):
    m, T, U = setup(dim)
    s, dt = 0.07, 0.1
    X = m.X.copy()
    zero = np.zeros_like(X)
    Ms, _, sp = assemble(m, T, None, kappa=None, adv=False, lump=lump)
    A1, r1, _ = assemble(m, T, None, kappa=None, adv=False, lump=lump, X_old=X, X_new=(1 + s) * X, ug=zero,
                         move=True, dt=dt)
    f = (1 + s) ** dim
    assert rel(A1, f * Ms) < 1e-13
    assert np.allclose(r1, -(f - 1) * csr_matvec(sp, Ms, T) / dt, rtol=1e-12, atol=1e-14 * np.abs(r1).max())


@pytest.mark.parametrize("dim,ibp", [(2, False), (3, False), (2, True), (3, True)])
def test_translation_is_relative_velocity(dim, ibp):
    m, T, U = setup(dim)
    dt = 0.1
    c = np.array([0.3, -0.2, 0.1])[:dim]
    X = m.X.copy()
    ug = np.tile(c / dt, (m.nx, 1))
    A_s, r_s, _ = assemble(m, T, U - c / dt, ibp=ibp, beta=0.0, dt=dt)
    A_m, r_m, _ = assemble(m, T, U, ibp=ibp, beta=0.0, X_old=X, X_new=X + c, ug=ug, move=True, dt=dt)
    assert rel(A_m, A_s) < 1e-12 and rel(r_m, r_s) < 1e-11


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_divergence_term(dim):
    m, T, _ = setup(dim)
    G = np.array([[0.3, -0.7, 0.2], [1.1, 0.4, -0.5], [0.2, 0.9, 0.6]])[:dim, :dim]
    ug = m.X @ G.T       # linear: div = trace(G); the flow moves with the grid
    X = m.X.copy()
    dt = 0.1
    Ms, _, sp = assemble(m, T, None, kappa=None, adv=False, dt=dt)
    A, r, _ = assemble(m, T, ug, kappa=None, ibp=False, X_old=X, X_new=X, ug=ug, move=True, dt=dt)
    div = np.trace(G)
    assert rel(A, Ms - dt * 0.5 * div * Ms) < 1e-12
    assert np.allclose(r, div * csr_matvec(sp, Ms, T), rtol=1e-11, atol=1e-13 * np.abs(r).max())
    # by parts: no divergence term, the relative velocity is zero
    A_i, r_i, _ = assemble(m, T, ug, kappa=None, ibp=True, beta=1.0, X_old=X, X_new=X, ug=ug, move=True, dt=dt)
    assert rel(A_i, Ms) < 1e-12 and np.abs(r_i).max() < 1e-12 * np.abs(r).max()


def test_motion_refuses_what_the_reference_refuses():
    m, T, U = setup(2)
    X = m.X.copy()
    el = E.make_element(2, 1, 2)
    fields = {"T": T, "U": U, "kappa": None, "src": None, "absorp": None, "X_old": X, "X_new": X,
              "ug": np.zeros_like(X)}
    sp = O.make_sparsity(m)
    col, ncol = O.colour_elements(m)
    asm = O.Assembly(m, el, el, fields, sp, O.colour_sets(col, ncol))
    with pytest.raises(ValueError):
        O.assemble(O.params(0.1, 0.5, 0.0, diff=False, stab="supg", move_mesh=True), asm)

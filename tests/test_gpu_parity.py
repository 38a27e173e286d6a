"""GPU parity of the gfx950 library against the oracle (CPU restatement of
Fluidity's CG advection-diffusion path), through the C ABI.

Tolerances: values are compared with the norm-wise relative error
max|A_gpu - A_ref| / max|A_ref| (and the same for the rhs); 1e-12 for
FP64 element arithmetic (the kernels use fused multiply-adds and
contracted quadrature tables, the oracle follows the Fortran source order),
1e-5 for FP32 element arithmetic.  The sparsity pattern is compared bit for
bit.
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, meshgen as M

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
FP32_RTOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_case(cfg, opts_kw=None, precision="fp64", scheme="auto", isotropic=True, kappa_full=None):
    print(f"\n[case] {cfg.name} n={cfg.mesh.nele} precision={precision} scheme={scheme} {opts_kw}", flush=True)
    ref = O.reference_system(cfg) if kappa_full is None else None
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg, precision=precision, scheme=scheme)
    for k, v in (opts_kw or {}).items():
        setattr(opts, k, v)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, inact = pr.matrix()
    return pr, ref, A, rhs, inact


@pytest.mark.parametrize("name,scale,scheme", [("c2", 6, "colour"), ("c2", 6, "blocked"), ("c1", 16, "colour"),
                                               ("c1", 16, "blocked"), ("c3", 3, "colour"), ("c4", 6, "blocked"),
                                               ("c5", 6, "blocked"), ("c5", 9, "colour"), ("c1p2", 6, "colour"),
                                               ("cavity", None, "colour"), ("cavity", None, "blocked"),
                                               ("cube1", None, "colour"), ("cube1", None, "blocked"),
                                               ("lapgrid", None, "colour"), ("lapgrid", None, "blocked")])
def test_assembly_matches_oracle(require_gpu, name, scale, scheme):
    cfg = M.config(name, seed=3, scale=scale)
    pr, ref, A, rhs, inact = run_case(cfg, scheme=scheme)
    f, c, _ = ref["sparsity"]
    assert np.array_equal(pr.sparsity.findrm, f) and np.array_equal(pr.sparsity.colm, c)
    assert np.array_equal(inact, ref["inactive"])
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL
    pr.close()


@pytest.mark.parametrize("variant", [
    dict(lump_mass_matrix=True),
    dict(integrate_advection_by_parts=True),
    dict(beta=1.0),
    dict(integrate_advection_by_parts=True, beta=0.3),
    dict(have_absorption=True),
])
def test_assembly_option_variants(require_gpu, variant):
    cfg = M.config("c1", seed=5, scale=12)
    # give every term something to do
    cfg.source = np.cos(3 * cfg.mesh.node_pos[:, 0])
    cfg.absorption = 0.5 + cfg.mesh.node_pos[:, 1] if variant.get("have_absorption") else None
    ref = O.reference_system_opts(cfg, variant)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg)
    for k, v in variant.items():
        setattr(opts, k, v)
    print(f"\n[case] c1 variant {variant}", flush=True)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL
    pr.close()


def test_assembly_p2_fp32(require_gpu):
    cfg = M.config("c3", seed=2, scale=3)
    pr, ref, A, rhs, _ = run_case(cfg, precision="fp32")
    assert rel(A, ref["A"]) < FP32_RTOL
    assert rel(rhs, ref["rhs"]) < FP32_RTOL
    pr.close()


@pytest.mark.parametrize("name,scale", [("c1", 20), ("c2", 10), ("c5", 12), ("cavity", None)])
def test_blocked_equals_colour_bitwise(require_gpu, name, scale):
    """Both schemes add each entry's element contributions in colour order:
    the results are identical bit for bit."""
    cfg = M.config(name, seed=13, scale=scale)
    pr = cg.problem_from_config(cfg)
    out = {}
    for scheme in ("colour", "blocked"):
        print(f"\n[case] {name} scale {scale} scheme {scheme}", flush=True)
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, scheme=scheme), cfg.dt)
        out[scheme] = pr.matrix()
    assert np.array_equal(out["colour"][0], out["blocked"][0])
    assert np.array_equal(out["colour"][1], out["blocked"][1])
    pr.close()


@pytest.mark.parametrize("name,scale", [("c1", 20), ("c5", 12), ("cavity", None)])
def test_blocked_fp32_equals_colour_fp32(require_gpu, name, scale):
    """FP32 element arithmetic: the node-staged slab kernel and the colour
    scheme add the same FP32 element values, widened, into the FP64 matrix
    in the same order -- identical bits; and both stay within the FP32
    bound of the FP64 assembly."""
    cfg = M.config(name, seed=15, scale=scale)
    pr = cg.problem_from_config(cfg)
    out = {}
    for scheme, prec in (("colour", "fp32"), ("blocked", "fp32"), ("blocked", "fp64")):
        print(f"\n[case] {name} scale {scale} scheme {scheme} {prec}", flush=True)
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, precision=prec, scheme=scheme), cfg.dt)
        out[(scheme, prec)] = pr.matrix()
    assert np.array_equal(out[("colour", "fp32")][0], out[("blocked", "fp32")][0])
    assert np.array_equal(out[("colour", "fp32")][1], out[("blocked", "fp32")][1])
    assert rel(out[("blocked", "fp32")][0], out[("blocked", "fp64")][0]) < FP32_RTOL
    pr.close()


def test_auto_scheme_falls_back_to_colour(require_gpu):
    """A fan of 400 triangles around one node: its row (401 entries) and its
    400 elements exceed a slab, so the automatic scheme is the colour scheme
    (and asking for the blocked one is an error); the result matches the
    oracle."""
    n = 400
    ang = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    X = np.vstack([[0.0, 0.0], np.stack([np.cos(ang), np.sin(ang)], axis=1)])
    conn = (np.array([[0, 1 + k, 1 + (k + 1) % n] for k in range(n)]) + 1).astype(np.int32).ravel()
    mesh = M.Mesh(2, n, n + 1, conn, X, 1, n + 1, conn, X, np.arange(2, n + 2, dtype=np.int32), name="fan")
    cfg = M.config("c1", seed=3, scale=4)
    cfg.mesh = mesh
    cfg.T = np.cos(2.0 * X[:, 0]) + X[:, 1]
    cfg.U = np.stack([-X[:, 1], X[:, 0]], axis=1)
    ref = O.reference_system(cfg)
    pr = cg.problem_from_config(cfg)
    from fluidity_amd.lib import FlgpuError
    with pytest.raises(FlgpuError):
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, scheme="blocked"), cfg.dt)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL
    # FP32 element arithmetic: the colour scheme, named or chosen, needs no plan
    for scheme in ("colour", "auto"):
        print(f"\n[case] fan fp32 {scheme}", flush=True)
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, precision="fp32", scheme=scheme), cfg.dt)
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        A32, rhs32, _ = pr.matrix()
        assert rel(A32, ref["A"]) < FP32_RTOL
        assert rel(rhs32, ref["rhs"]) < FP32_RTOL
    # SUPG: the automatic scheme gives way to the colour kernel as well
    print("\n[case] fan supg", flush=True)
    var = {"stabilisation": "supg"}
    ref_s = O.reference_system_opts(cfg, var)
    for scheme in ("blocked", "auto"):
        opts = cg.options_from_config(cfg, scheme=scheme)
        opts.stabilisation = "supg"
        if scheme == "blocked":
            with pytest.raises(FlgpuError):
                cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
            continue
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
        cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
        A_s, rhs_s, _ = pr.matrix()
        assert rel(A_s, ref_s["A"]) < FP64_RTOL
        assert rel(rhs_s, ref_s["rhs"]) < FP64_RTOL
    pr.close()


@pytest.mark.parametrize("name,scale", [("c2", 8), ("c1", 12)])
def test_blocked_without_node_staging_equals_colour(require_gpu, name, scale, monkeypatch):
    """The blocked scheme's gather form (used when the field mesh is not the
    coordinate mesh or the diffusivity is anisotropic) gives the same bits."""
    monkeypatch.setenv("FLGPU_NO_STAGING", "1")
    cfg = M.config(name, seed=14, scale=scale)
    pr = cg.problem_from_config(cfg)
    out = {}
    for scheme in ("colour", "blocked"):
        print(f"\n[case] {name} scale {scale} scheme {scheme} (no staging)", flush=True)
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, scheme=scheme), cfg.dt)
        out[scheme] = pr.matrix()
    assert np.array_equal(out["colour"][0], out["blocked"][0])
    assert np.array_equal(out["colour"][1], out["blocked"][1])
    pr.close()


def test_assembly_deterministic(require_gpu):
    cfg = M.config("c2", seed=7, scale=8)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg)
    outs = []
    for _ in range(3):
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
        outs.append(pr.matrix())
    for a, b in zip(outs[1:], outs[:-1]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    pr.close()


def test_pcg_fp64_matches_oracle(require_gpu):
    cfg = M.config("c2", seed=3, scale=8)
    pr, ref, A, rhs, inact = run_case(cfg)
    so = cg.SolverOptions("cg", "jacobi", relative_error=1e-10, norm="unpreconditioned")
    x, info = cg.petsc_solve(pr, so)
    xr, its, reason, relres, _ = O.petsc_solve("cg", "jacobi", 1e-12, ref["sparsity"], ref["A"], ref["inactive"],
                                               ref["rhs"])
    print("gpu its", info.its, "reason", info.reason, "relres", info.true_relres, "oracle its", its, relres)
    assert info.reason > 0 and reason > 0
    assert info.true_relres <= 1e-10
    assert rel(x, xr) < 1e-8
    # ghost rows take the rhs values (Petsc_Tools.F90:724-731)
    b = cfg.mesh.boundary - 1
    assert np.array_equal(x[b], rhs[b])
    pr.close()


def test_pcg_mixed_precision(require_gpu):
    cfg = M.config("c4", seed=4, scale=10)
    pr, ref, A, rhs, inact = run_case(cfg)
    so64 = cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned")
    x64, i64 = cg.petsc_solve(pr, so64)
    so32 = cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned", precision="mixed")
    x32, i32 = cg.petsc_solve(pr, so32)
    print("fp64", i64.its, i64.true_relres, "mixed", i32.its, i32.true_relres, "refinements", i32.refinements)
    assert i32.reason > 0 and i32.true_relres <= 1e-8
    assert rel(x32, x64) < 1e-6
    pr.close()


def test_gmres_c1_matches_scipy(require_gpu):
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl

    cfg = M.config("c1", seed=1, scale=16)
    pr, ref, A, rhs, inact = run_case(cfg)
    so = cg.SolverOptions("gmres", "jacobi", relative_error=1e-10, restart=30)
    x, info = cg.petsc_solve(pr, so)
    f, c, cen = ref["sparsity"]
    n = f.size - 1
    S = sps.csr_matrix((ref["A"], c - 1, f - 1), shape=(n, n)).tolil()
    b = ref["rhs"].copy()
    ghost = ref["inactive"].astype(bool)
    # reduced system: lifting + ghost pivot
    g = np.where(ghost, b, 0.0)
    bb = b - sps.csr_matrix(S) @ g
    S = S.tocsr()
    S[:, ghost] = 0.0
    S[ghost, :] = 0.0
    piv = O.ghost_pivot(cen, ref["A"])
    S = S.tolil()
    for i in np.nonzero(ghost)[0]:
        S[i, i] = piv
    bb[ghost] = 0.0
    xs = spl.spsolve(S.tocsr(), bb)
    xs[ghost] = b[ghost]
    print("gmres its", info.its, "reason", info.reason, "relres", info.true_relres)
    assert info.reason > 0 and info.true_relres < 1e-9
    assert rel(x, xs) < 1e-8
    pr.close()


def test_solve_field_equation_cg_step(require_gpu):
    cfg = M.config("c1", seed=2, scale=16)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg)
    so = cg.SolverOptions("gmres", "jacobi", relative_error=1e-10)
    T1, its, info = cg.solve_field_equation_cg(pr, opts, so, cfg.dt, cfg.mesh.boundary, 0.0)
    assert info.reason > 0 and its > 0
    b = cfg.mesh.boundary - 1
    np.testing.assert_allclose(T1[b], 0.0, atol=1e-14)  # Dirichlet 0 reached in one step
    # the bump is advected and diffused but stays bounded
    assert np.max(T1) <= np.max(cfg.T) * 1.05
    pr.close()


def test_out_of_pattern_insertion_is_an_error(require_gpu):
    cfg = M.config("c2", seed=1, scale=3)
    sp = cg.make_sparsity(cfg.mesh)
    # drop one off-diagonal entry of row 1 from the pattern
    f = sp.findrm.copy()
    c = sp.colm.copy()
    drop = 1 if c[0] == 1 else 0  # an entry that is not the diagonal
    keep = np.ones(c.size, bool)
    keep[drop] = False
    f[1:] -= 1
    bad = cg.Sparsity(f, c[keep].copy(), np.where(sp.centrm > drop + 1, sp.centrm - 1, sp.centrm).astype(np.int32))
    pr = cg.CGProblem(cfg.mesh, quad_degree=2, sparsity=bad)
    pr.set_fields(T=cfg.T, diffusivity=1.0, source=cfg.source)
    opts = cg.options_from_config(cfg)
    from fluidity_amd.lib import FlgpuError
    with pytest.raises(FlgpuError, match="outside sparsity"):
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    pr.close()


@pytest.mark.parametrize("n", [20000, 70000])
def test_csr_mult_bitwise_matches_oracle(require_gpu, n):
    """The streamed CSR product sums each row in entry order with the product
    rounded first, as csr_mult does (femtools/Sparse_Tools.F90:3822-3844):
    the result equals the oracle bit for bit.  The pattern mixes short rows,
    empty rows and one row longer than a staged block (the chunked path);
    with n = 20000 every column lies within int16 range of its row block
    (16-bit column offsets), with n = 70000 it does not (int32 columns)."""
    import ctypes as C
    from fluidity_amd import lib as L

    rng = np.random.default_rng(5)
    lens = rng.integers(1, 30, n)
    lens[17] = 0
    lens[4321] = 5000             # longer than one staged block of entries
    lens[4322] = 2048             # exactly one block
    rows = []
    for i, m in enumerate(lens):
        cols = np.unique(np.concatenate([[i], rng.integers(0, n, m)]))[: max(m, 1)] if m else np.array([], int)
        rows.append(np.sort(cols))
    findrm = np.concatenate([[1], 1 + np.cumsum([r.size for r in rows])]).astype(np.int32)
    colm = (np.concatenate(rows) + 1).astype(np.int32)
    centrm = np.zeros(n, np.int32)
    for i, r in enumerate(rows):
        hit = np.nonzero(r == i)[0]
        if hit.size:
            centrm[i] = findrm[i] + hit[0]
    val = rng.standard_normal(colm.size)
    val[::7] = 0.0
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    lib = L.load()
    ctx = C.c_void_p()
    L.check(lib.flgpu_init(C.byref(ctx), 0), "init")
    try:
        L.check(lib.flgpu_sparsity_bind(ctx, n, n, L.ptr(findrm), L.ptr(colm), L.ptr(centrm)), "sparsity_bind")
        L.check(lib.flgpu_matrix_set(ctx, L.ptr(val), L.ptr(np.zeros(n)), None), "matrix_set")
        y = np.zeros(n)
        L.check(lib.flgpu_csr_mult(ctx, L.ptr(x), L.ptr(y)), "csr_mult")
    finally:
        lib.flgpu_finalize(ctx)
    yr = O.csr_mult(findrm, colm, val, x)
    assert np.array_equal(y, yr)

"""GPU parity of the face loop (flgpu_surface_bind / flgpu_advdif_surface,
the face loop of assemble_advection_diffusion_cg,
assemble/Advection_Diffusion_CG.F90:606-638, :1228-1379) against the
oracle, through the C ABI.

Tolerance: the element assembly's, max|A_gpu - A_ref| / max|A_ref| < 1e-12
(FP64); the face kernel follows the oracle's operation order without
contraction, so the face terms themselves usually agree bit for bit."""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, lib as L, meshgen as M

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run(cfg, kind, variant, scheme="auto"):
    print(f"\n[case] {cfg.name} n={cfg.mesh.nele} bcs={kind} {variant} scheme={scheme}", flush=True)
    ref = O.reference_system(cfg, variant=dict(variant, bcs=kind))
    pr = cg.problem_from_config(cfg)
    facets = cfg.mesh.meta.get("surface_facets")  # a mesh file's own surface elements, in file order
    surf = M.surface_from_facets(cfg.mesh, facets) if facets else M.surface_of(cfg.mesh)
    pr.bind_surface(surf)
    bt, b1, b2, strong = M.unit_box_bcs(cfg.mesh, surf, kind)
    opts = cg.options_from_config(cfg, scheme=scheme)
    for k, v in variant.items():
        setattr(opts, k, v)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt, cg.SurfaceBCs(bt, b1, b2))
    cg.apply_dirichlet_conditions(pr, strong, 0.0, cfg.dt)
    A, rhs, inact = pr.matrix()
    pr.close()
    return ref, A, rhs, inact


@pytest.mark.parametrize("name,scale,variant,scheme", [
    ("c1", 12, {}, "colour"),
    ("c1", 12, {"integrate_advection_by_parts": True}, "blocked"),
    ("c2", 5, {}, "auto"),
    ("c5", 5, {}, "blocked"),
    ("c1p2", 5, {"integrate_advection_by_parts": True, "beta": 0.3}, "colour"),
    ("c3", 2, {"integrate_advection_by_parts": True}, "colour"),
    ("cube1", None, {"integrate_advection_by_parts": True}, "colour"),
    ("lapgrid", None, {}, "blocked"),
])
def test_surface_matches_oracle(require_gpu, name, scale, variant, scheme):
    cfg = M.config(name, seed=7, scale=scale)
    ref, A, rhs, inact = run(cfg, "neumann_robin", variant, scheme)
    assert np.array_equal(inact, ref["inactive"])
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL


def test_weak_dirichlet_inflow_matches_oracle(require_gpu):
    cfg = M.config("c1", seed=8, scale=10)
    cfg.U = np.tile(np.array([1.0, 0.25]), (cfg.mesh.nnodes, 1))
    cfg.kappa = None
    ref, A, rhs, inact = run(cfg, "weak", {"integrate_advection_by_parts": True, "have_diffusivity": False})
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL


def test_weak_dirichlet_with_diffusivity_is_an_error(require_gpu):
    cfg = M.config("c1", seed=8, scale=6)
    pr = cg.problem_from_config(cfg)
    surf = M.surface_of(cfg.mesh)
    pr.bind_surface(surf)
    bt, b1, b2, _ = M.unit_box_bcs(cfg.mesh, surf, "weak")
    with pytest.raises(L.FlgpuError) as e:
        cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt, cg.SurfaceBCs(bt, b1, b2))
    assert e.value.code == 6
    pr.close()


def test_neumann_robin_linear_solution_on_gpu(require_gpu):
    """kappa dT/dn = g0 on x = 0, = g1 - h T on x = 1, no flux elsewhere:
    the exact solution a + b x (b = -g0, a = (g1 - b)/h - b) is in P1, so
    the GPU assembly + CG solve reproduces it at the nodes."""
    cfg = M.config("c2", seed=9, scale=8)
    cfg.kappa = 1.0
    cfg.source = None
    pr = cg.problem_from_config(cfg)
    surf = M.surface_of(cfg.mesh)
    pr.bind_surface(surf)
    bt, _, _, _ = M.unit_box_bcs(cfg.mesh, surf, "neumann_robin")
    g0, g1, h = 0.8, 1.5, 2.5
    b1 = np.where(np.repeat(bt == M.BC_NEUMANN, surf.floc), g0, g1)
    b2 = np.full(surf.nfaces * surf.floc, h)
    opts = cg.CGOptions(theta=1.0, exclude_mass_terms=True, exclude_advection_terms=True)
    solver = cg.SolverOptions(iterative_method="cg", preconditioner="jacobi", relative_error=1e-12,
                              norm="unpreconditioned")
    T1, its, info = cg.solve_field_equation_cg(pr, opts, solver, 1.0, bcs=cg.SurfaceBCs(bt, b1, b2))
    assert info.reason > 0
    b = -g0
    a = (g1 - b) / h - b
    np.testing.assert_allclose(T1, a + b * cfg.mesh.node_pos[:, 0], rtol=0, atol=1e-8)
    pr.close()


def test_surface_lists_follow_a_sparsity_rebind(require_gpu):
    """The face addition lists are rebuilt when the sparsity is rebound;
    two calls in a row give the same bits."""
    cfg = M.config("c2", seed=10, scale=4)
    pr = cg.problem_from_config(cfg)
    surf = M.surface_of(cfg.mesh)
    pr.bind_surface(surf)
    bt, b1, b2, _ = M.unit_box_bcs(cfg.mesh, surf, "neumann_robin")
    opts = cg.options_from_config(cfg)
    out = []
    for k in range(3):
        if k == 2:
            pr.bind(cfg.mesh, quad_degree=cfg.quad_degree)  # rebinding the field mesh unbinds the surface
            pr.set_fields(T=cfg.T, diffusivity=cfg.kappa, source=cfg.source)
            with pytest.raises(L.FlgpuError):
                cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt, cg.SurfaceBCs(bt, b1, b2))
            pr.bind_surface(surf)
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt, cg.SurfaceBCs(bt, b1, b2))
        out.append(pr.matrix()[:2])
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
    pr.close()


@pytest.mark.parametrize("name,scale,kind,opts", [
    ("c1", 10, "neumann_robin", dict(equation_type="internal_energy", density_theta=0.6)),
    ("c2", 4, "neumann_robin", dict(equation_type="internal_energy", density_theta=0.3, multiphase=True,
                                    compressible=True)),
    ("c1", 10, "weak", dict(equation_type="internal_energy", multiphase=True, have_diffusivity=False)),
    ("c1", 10, "neumann_robin", dict(move_mesh=True)),
    ("c2", 4, "neumann_robin", dict(move_mesh=True, beta=0.4)),
])
def test_surface_equation_types_and_moving_mesh(require_gpu, name, scale, kind, opts):
    """The face advection by parts with the internal energy's density (and
    volume fraction) and with the velocity relative to the grid."""
    cfg = M.config(name, seed=9, scale=scale)
    p = cfg.mesh.node_pos
    if kind == "weak":
        cfg.U = np.tile(np.array([1.0, 0.25]), (cfg.mesh.nnodes, 1))
        cfg.kappa = None
    if cfg.U is None:
        cfg.U = M.abc_flow(p)
        cfg.have_advection = cfg.have_mass = True
        cfg.dt, cfg.theta = 0.005, 0.5
    fields = dict(density=1.0 + 0.5 * p[:, 0], old_density=0.7 + 0.3 * p[:, 1], volume_fraction=0.4 + 0.2 * p[:, 0],
                  pressure=1.5 + p[:, 1])
    X = cfg.mesh.X
    motion = dict(old_positions=X - 0.01 * np.sin(np.pi * X), new_positions=X + 0.01 * np.sin(2 * np.pi * X))
    motion["grid_velocity"] = (motion["new_positions"] - motion["old_positions"]) / cfg.dt
    opts = dict(opts, integrate_advection_by_parts=True)
    variant = dict(opts, bcs=kind)
    variant.update(fields if opts.get("equation_type") else motion)
    print(f"\n[case] {name} n={cfg.mesh.nele} bcs={kind} {opts}", flush=True)
    ref = O.reference_system(cfg, variant=variant)
    pr = cg.problem_from_config(cfg)
    if opts.get("equation_type"):
        pr.set_fields(**fields)
    else:
        pr.set_mesh_motion(motion["old_positions"], motion["new_positions"], motion["grid_velocity"])
    surf = M.surface_of(cfg.mesh)
    pr.bind_surface(surf)
    bt, b1, b2, strong = M.unit_box_bcs(cfg.mesh, surf, kind)
    o = cg.options_from_config(cfg)
    for k, v in opts.items():
        setattr(o, k, v)
    cg.assemble_advection_diffusion_cg(pr, o, cfg.dt, cg.SurfaceBCs(bt, b1, b2))
    cg.apply_dirichlet_conditions(pr, strong, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    pr.close()
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL

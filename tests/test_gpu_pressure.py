"""The pressure path on the GPU (flgpu_pressure_bind / flgpu_ct_assemble /
flgpu_cmc_assemble) against the oracle, through the C ABI.

The GPU kernels follow the oracle's operation order without contraction,
so ct_m, the lumped mass and C^T M_l^-1 C are compared bit for bit: on
P1-P1 and Taylor-Hood P2-P1 meshes in 2-D and 3-D, by parts and not, with
a constant and a nodal density, with a given inverse lumped mass and with
the device's own 1/masslump; at the full size of the 3-D P1 configuration
(64^3 x 6 tets); and an element entry outside ct_m's pattern is an error.
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, meshgen as M
from fluidity_amd.lib import FlgpuError

pytestmark = pytest.mark.gpu

MESHES = [("p1p1-2d", lambda: M.square_tris(12, jitter=0.2, seed=3)),
          ("p1p1-3d", lambda: M.cube_tets(8, jitter=0.15, seed=4)),
          ("p2p1-2d", lambda: M.square_tris(8, jitter=0.2, seed=5, degree=2)),
          ("p2p1-3d", lambda: M.cube_tets(5, jitter=0.15, seed=6, degree=2))]


def quad_degree(mesh):
    return 2 if mesh.degree == 1 else 4


def gpu_problem(mesh):
    pr = cg.CGProblem(mesh, quad_degree=quad_degree(mesh))
    pm = cg.vertex_pressure_mesh(pr)
    cg.bind_pressure_mesh(pr, pm)
    return pr, pm


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
@pytest.mark.parametrize("ibp", [False, True], ids=["grad", "by_parts"])
def test_ct_and_cmc_match_oracle_bitwise(require_gpu, name, make, ibp):
    mesh = make()
    print(f"\n[case] pressure {name} ibp={ibp} nele={mesh.nele}", flush=True)
    pr, pm = gpu_problem(mesh)
    col = (pr.colouring[1] - 1, pr.colouring[2])  # the oracle takes 0-based set offsets
    for density in (1.0, 1.0 + 0.5 * mesh.node_pos[:, 0]):
        asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, quad_degree(mesh), colouring=col, ibp=ibp,
                                                     density=density)
        assert np.array_equal(cts[0], pm.ct_findrm) and np.array_equal(cts[1], pm.ct_colm)
        ct_o, ml_o = O.ct_assemble(asm, pd)
        ct_g, ml_g = cg.construct_momentum_ct(pr, pm, integrate_continuity_by_parts=ibp, density=density)
        assert np.array_equal(ct_g, ct_o), np.max(np.abs(ct_g - ct_o))
        assert np.array_equal(ml_g, ml_o), np.max(np.abs(ml_g - ml_o))
    cf, cc = cg.cmc_sparsity(pm, mesh.nnodes)
    inv = np.random.default_rng(4).uniform(0.5, 2.0, size=(mesh.dim, mesh.nnodes))
    inv[:, mesh.boundary - 1] = 0.0  # Dirichlet velocity rows zeroed, as apply_dirichlet_conditions_inverse_mass
    cmc_g = cg.assemble_masslumped_cmc(pr, cf, cc, inv)
    cmc_o = O.cmc(cts, ct_g, inv, (cf, cc))
    assert np.array_equal(cmc_g, cmc_o), np.max(np.abs(cmc_g - cmc_o))
    # the device's own inverse of the lumped mass (every component 1/ml)
    cmc_g1 = cg.assemble_masslumped_cmc(pr, cf, cc, None)
    cmc_o1 = O.cmc(cts, ct_g, np.tile(1.0 / ml_g, (mesh.dim, 1)), (cf, cc))
    assert np.array_equal(cmc_g1, cmc_o1)
    pr.close()


def test_full_size_p1_configuration(require_gpu):
    """64^3 x 6 tets (the 3-D P1 configuration's mesh): ct_m, the lumped
    mass and C^T M_l^-1 C equal the oracle's bit for bit; the discrete
    divergence of a linear field is exact."""
    cfg = M.config("c2", seed=1)
    mesh = cfg.mesh
    print(f"\n[case] pressure full size c2 nele={mesh.nele}", flush=True)
    pr, pm = gpu_problem(mesh)
    ct_g, ml_g = cg.construct_momentum_ct(pr, pm)
    asm, pd, cts, pn, p_nodes = O.pressure_setup(mesh, 2, colouring=(pr.colouring[1] - 1, pr.colouring[2]))
    ct_o, ml_o = O.ct_assemble(asm, pd)
    assert np.array_equal(ct_g, ct_o) and np.array_equal(ml_g, ml_o)
    cf, cc = cg.cmc_sparsity(pm, mesh.nnodes)
    cmc_g = cg.assemble_masslumped_cmc(pr, cf, cc, None)
    cmc_o = O.cmc(cts, ct_o, np.tile(1.0 / ml_o, (mesh.dim, 1)), (cf, cc))
    assert np.array_equal(cmc_g, cmc_o)
    u = mesh.node_pos @ np.array([[1.0, 2.0, 0.0], [0.5, -3.0, 1.0], [0.0, 0.2, 0.5]]).T
    rows = np.repeat(np.arange(p_nodes), np.diff(pm.ct_findrm))
    div = np.zeros(p_nodes)
    for a in range(mesh.dim):
        np.add.at(div, rows, ct_g[a] * u[pm.ct_colm - 1, a])
    # the P1 pressure basis integrates to the lumped P1 mass with unit density
    assert np.max(np.abs(div - (1.0 - 3.0 + 0.5) * ml_g)) <= 1e-12 * np.max(np.abs(div))
    pr.close()


def test_time_pressure_leaves_results(require_gpu):
    """flgpu_time_pressure re-runs the kernels on the device-resident inputs:
    device times are positive and ct_m / C^T M_l^-1 C (own inverse) read
    back afterwards equal the first results bit for bit."""
    import ctypes as C
    from fluidity_amd import lib as L
    mesh = M.cube_tets(6, jitter=0.15, seed=8)
    pr, pm = gpu_problem(mesh)
    ct1, ml1 = cg.construct_momentum_ct(pr, pm)
    cf, cc = cg.cmc_sparsity(pm, mesh.nnodes)
    cmc1 = cg.assemble_masslumped_cmc(pr, cf, cc, None)
    a, b = C.c_double(0), C.c_double(0)
    with pytest.raises(FlgpuError):
        L.check(pr.lib.flgpu_time_pressure(pr.ctx, 0, C.byref(a), C.byref(b)), "time_pressure")
    L.check(pr.lib.flgpu_time_pressure(pr.ctx, 3, C.byref(a), C.byref(b)), "time_pressure")
    assert a.value > 0 and b.value > 0
    cmc2 = cg.assemble_masslumped_cmc(pr, cf, cc, None)
    ct2, ml2 = cg.construct_momentum_ct(pr, pm)
    assert np.array_equal(cmc1, cmc2) and np.array_equal(ct1, ct2) and np.array_equal(ml1, ml2)
    pr.close()


def test_ct_out_of_pattern_is_an_error(require_gpu):
    mesh = M.cube_tets(3, jitter=0.1, seed=7)
    pr = cg.CGProblem(mesh, quad_degree=2)
    pm = cg.vertex_pressure_mesh(pr)
    f, c = pm.ct_findrm, pm.ct_colm
    keep = np.ones(c.size, bool)
    keep[f[1] - 2] = False  # the last entry of row 1
    f2 = f.copy()
    f2[1:] -= 1
    bad = cg.PressureMesh(pm.ndglno, pm.loc, pm.nnodes, pm.element, f2, c[keep].copy())
    cg.bind_pressure_mesh(pr, bad)
    print("\n[case] pressure out of pattern", flush=True)
    with pytest.raises(FlgpuError, match="outside sparsity"):
        cg.construct_momentum_ct(pr, bad)
    with pytest.raises(FlgpuError, match="assemble ct_m first"):
        cg.assemble_masslumped_cmc(pr, *cg.cmc_sparsity(pm, mesh.nnodes))
    pr.close()

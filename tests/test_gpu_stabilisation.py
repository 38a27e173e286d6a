"""GPU parity of the stabilised assembly (streamline upwind and SUPG,
assemble/Advection_Diffusion_CG.F90:770-825, :1110-1121,
assemble/Upwind_Stabilisation.F90) against the oracle, through the C ABI.
Tolerances as for the element assembly: 1e-12 relative (FP64 element
arithmetic), 1e-5 (FP32)."""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, lib as L, meshgen as M

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
FP32_RTOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run(cfg, variant, precision="fp64"):
    print(f"\n[case] {cfg.name} n={cfg.mesh.nele} {variant} {precision}", flush=True)
    ref = O.reference_system_opts(cfg, variant)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg, precision=precision)
    for k, v in variant.items():
        setattr(opts, k, v)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    pr.close()
    return ref, A, rhs


@pytest.mark.parametrize("name,scale,variant", [
    ("c1", 12, {"stabilisation": "supg"}),
    ("c1", 12, {"stabilisation": "streamline_upwind"}),
    ("c1", 12, {"stabilisation": "supg", "integrate_advection_by_parts": True, "beta": 0.3}),
    ("c1", 12, {"stabilisation": "supg", "nu_bar_scheme": "critical_rule", "lump_mass_matrix": True}),
    ("c1", 12, {"stabilisation": "streamline_upwind", "nu_bar_scheme": "unity", "nu_scale": 1.0}),
    ("c1", 12, {"stabilisation": "supg", "nu_bar_scheme": "doubly_asymptotic", "have_absorption": True}),
    ("cavity", None, {"stabilisation": "supg"}),
    ("c1p2", 5, {"stabilisation": "supg", "beta": 0.4}),
    ("c3", 2, {"stabilisation": "supg"}),
    ("c3", 2, {"stabilisation": "streamline_upwind", "nu_bar_scheme": "doubly_asymptotic"}),
])
def test_stabilised_assembly_matches_oracle(require_gpu, name, scale, variant):
    cfg = M.config(name, seed=11, scale=scale)
    cfg.source = np.cos(3 * cfg.mesh.node_pos[:, 0])
    if variant.get("have_absorption"):
        cfg.absorption = 0.5 + cfg.mesh.node_pos[:, 1]
    ref, A, rhs = run(cfg, variant)
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL


def test_stabilised_assembly_fp32(require_gpu):
    cfg = M.config("c3", seed=12, scale=2)
    ref, A, rhs = run(cfg, {"stabilisation": "supg"}, precision="fp32")
    assert rel(A, ref["A"]) < FP32_RTOL
    assert rel(rhs, ref["rhs"]) < FP32_RTOL


def test_stabilisation_argument_errors(require_gpu):
    cfg = M.config("c1", seed=13, scale=6)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg, scheme="blocked")
    opts.stabilisation = "supg"
    with pytest.raises(L.FlgpuError) as e:
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    assert e.value.code == 6  # the blocked scheme has no stabilised kernel
    pr.close()
    cfg2 = M.config("c2", seed=13, scale=4)  # no velocity field
    pr = cg.problem_from_config(cfg2)
    opts = cg.options_from_config(cfg2)
    opts.stabilisation = "streamline_upwind"
    with pytest.raises(L.FlgpuError) as e:
        cg.assemble_advection_diffusion_cg(pr, opts, cfg2.dt)
    assert e.value.code == 1
    pr.close()

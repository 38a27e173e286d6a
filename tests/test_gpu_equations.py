"""The internal-energy and k-epsilon equation types of the element routine
(assemble/Advection_Diffusion_CG.F90:484-552 options, :867-1226 terms) on
the GPU against the oracle, through the C ABI: density-weighted mass and
advection with the div(density u) split, the phase volume fraction
(multiphase), the compressible p div(u) term; P1 and P2, 1-D to 3-D, FP64
to 1e-12 and FP32 to 1e-5 relative (as the other assembly tests), also
with SUPG and streamline-upwind stabilisation.  The equation types take the
two-pass scheme by default (or the colour scheme by name, the same bits);
stabilised, the colour scheme (the blocked scheme is refused with
FLGPU_EUNSUPPORTED).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, lib as L, meshgen as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def eq_fields(cfg):
    p = cfg.mesh.node_pos
    x = p[:, 0]
    y = p[:, 1] if p.shape[1] > 1 else 0.5 * x
    return dict(density=1.0 + 0.5 * x + 0.3 * y ** 2, old_density=0.8 + 0.4 * y - 0.2 * x,
                pressure=2.0 + np.sin(3 * x), volume_fraction=0.3 + 0.4 * x * y)


def make_cfg(name, scale):
    cfg = M.config(name, seed=13, scale=scale)
    if cfg.U is None:  # give the 3-D pressure-Poisson mesh a velocity and mass
        p = cfg.mesh.node_pos
        cfg.U = M.abc_flow(p)
        cfg.T = np.sin(2 * p[:, 0]) * np.cos(p[:, 1] + p[:, 2])
        cfg.kappa = 1e-2
        cfg.dt, cfg.theta = 0.005, 0.5
        cfg.have_mass = cfg.have_advection = True
    cfg.source = np.cos(3 * cfg.mesh.node_pos[:, 0])
    return cfg


CASES = [
    ("c2", 5, dict(equation_type="internal_energy", density_theta=0.6, beta=0.4), "fp64"),
    ("c2", 5, dict(equation_type="internal_energy", density_theta=0.3, compressible=True, multiphase=True,
                   integrate_advection_by_parts=True, beta=0.3), "fp64"),
    ("c2", 5, dict(equation_type="k_epsilon", beta=1.0), "fp64"),
    ("c2", 5, dict(equation_type="k_epsilon", integrate_advection_by_parts=True, beta=0.0,
                   lump_mass_matrix=True), "fp64"),
    ("c1", 10, dict(equation_type="internal_energy", compressible=True, beta=1.0, lump_mass_matrix=True), "fp64"),
    ("c1", 10, dict(equation_type="internal_energy", multiphase=True, beta=0.5), "fp64"),
    ("c1p2", 4, dict(equation_type="internal_energy", compressible=True, multiphase=True, beta=0.7), "fp64"),
    ("c3", 2, dict(equation_type="k_epsilon", beta=0.5), "fp64"),
    ("interval", 64, dict(equation_type="internal_energy", compressible=True, multiphase=True, beta=0.2), "fp64"),
    ("c2", 5, dict(equation_type="internal_energy", compressible=True, multiphase=True, beta=0.4), "fp32"),
    ("c2", 5, dict(equation_type="internal_energy", density_theta=0.6, beta=0.4, stabilisation="supg",
                   compressible=True), "fp64"),
    ("c1", 10, dict(equation_type="internal_energy", multiphase=True, stabilisation="supg",
                    integrate_advection_by_parts=True, beta=0.3), "fp64"),
    ("c1p2", 4, dict(equation_type="k_epsilon", stabilisation="streamline_upwind", beta=1.0), "fp64"),
]


@pytest.mark.parametrize("name,scale,opts,precision", CASES,
                         ids=[f"{c[0]}-{c[2]['equation_type']}-{c[2].get('stabilisation', 'none')}-{i}-{c[3]}"
                              for i, c in enumerate(CASES)])
def test_equation_types_match_oracle(require_gpu, name, scale, opts, precision):
    cfg = make_cfg(name, scale)
    fields = eq_fields(cfg)
    print(f"\n[case] {name} n={cfg.mesh.nele} {opts} {precision}", flush=True)
    variant = dict(opts)
    variant.update(fields)
    ref = O.reference_system_opts(cfg, variant)
    pr = cg.problem_from_config(cfg)
    pr.set_fields(**fields)
    o = cg.options_from_config(cfg, precision=precision)
    for k, v in opts.items():
        setattr(o, k, v)
    cg.assemble_advection_diffusion_cg(pr, o, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    pr.close()
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel(A, ref["A"]) < tol
    assert rel(rhs, ref["rhs"]) < tol
    # the equation type changed the system (not the advection-diffusion one)
    base = O.reference_system_opts(cfg, {k: v for k, v in opts.items()
                                         if k in ("beta", "lump_mass_matrix", "integrate_advection_by_parts")})
    assert rel(base["A"], ref["A"]) > 1e-3 or rel(base["rhs"], ref["rhs"]) > 1e-3


@pytest.mark.parametrize("name,scale,opts", [
    ("c2", 4, dict(equation_type="internal_energy", compressible=True, multiphase=True, beta=0.4)),
    ("c1p2", 4, dict(equation_type="k_epsilon", integrate_advection_by_parts=True, beta=0.2)),
    ("interval", 32, dict(equation_type="internal_energy", lump_mass_matrix=True, beta=1.0)),
])
def test_two_pass_equals_colour_bitwise(require_gpu, name, scale, opts):
    """The default two-pass scheme (rows in one launch, colour-order sums
    per entry) gives the colour scheme's bits."""
    cfg = make_cfg(name, scale)
    pr = cg.problem_from_config(cfg)
    pr.set_fields(**eq_fields(cfg))
    mats = {}
    for scheme in ("auto", "colour", "blocked"):
        o = cg.options_from_config(cfg, scheme=scheme)
        for k, v in opts.items():
            setattr(o, k, v)
        cg.assemble_advection_diffusion_cg(pr, o, cfg.dt)
        mats[scheme] = pr.matrix()[:2]
    pr.close()
    for scheme in ("auto", "blocked"):
        assert np.array_equal(mats[scheme][0], mats["colour"][0]) and np.array_equal(mats[scheme][1], mats["colour"][1])


def test_equation_types_refuse_blocked_stabilised(require_gpu):
    cfg = make_cfg("c2", 3)
    pr = cg.problem_from_config(cfg)
    pr.set_fields(**eq_fields(cfg))
    for kw in (dict(scheme="blocked", stabilisation="supg"),):
        o = cg.options_from_config(cfg)
        o.equation_type = "internal_energy"
        for k, v in kw.items():
            setattr(o, k, v)
        p = o.params(cfg.dt)
        assert pr.lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)) == 6  # FLGPU_EUNSUPPORTED
    pr.close()


def test_equation_type_needs_its_fields(require_gpu):
    cfg = make_cfg("c2", 3)
    pr = cg.problem_from_config(cfg)
    o = cg.options_from_config(cfg)
    o.equation_type = "k_epsilon"
    p = o.params(cfg.dt)
    assert pr.lib.flgpu_advdif_assemble_resident(pr.ctx, C.byref(p)) == 1  # FLGPU_EINVAL: no density
    pr.close()


def test_equation_types_with_constant_fields(require_gpu):
    """Constant density, old density, pressure and volume fraction fields
    (FIELD_TYPE_CONSTANT: one value, interpolated like a nodal field)."""
    cfg = make_cfg("c1", 8)
    consts = dict(density=1.3, old_density=0.9, pressure=2.5, volume_fraction=0.6)
    opts = dict(equation_type="internal_energy", density_theta=0.4, compressible=True, multiphase=True, beta=0.5)
    print(f"\n[case] c1 constant fields {opts}", flush=True)
    ref = O.reference_system_opts(cfg, dict(opts, **{k: np.array([v]) for k, v in consts.items()}))
    pr = cg.problem_from_config(cfg)
    pr.set_fields(**consts)
    o = cg.options_from_config(cfg)
    for k, v in opts.items():
        setattr(o, k, v)
    cg.assemble_advection_diffusion_cg(pr, o, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    pr.close()
    assert rel(A, ref["A"]) < 1e-12
    assert rel(rhs, ref["rhs"]) < 1e-12

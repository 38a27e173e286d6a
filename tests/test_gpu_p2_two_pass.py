"""Quadratic T, the two-pass scheme (element values in one launch, then one
thread per CSR entry / rhs row adding them in colour order) against the
colour scheme and the oracle.

* Both schemes add each entry's element values in colour order, ascending
  element id inside a colour, so the matrices and right-hand sides are
  identical bit for bit (FP64 and FP32 element arithmetic, every term
  variant); the two-pass result is within 1e-12 (FP64) / 1e-5 (FP32) of the
  oracle's assembly.
* An element entry outside the sparsity pattern is an error, as in
  csr_addto.
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, meshgen as M
from fluidity_amd.lib import FlgpuError

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
FP32_RTOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


VARIANTS = [
    {},
    dict(lump_mass_matrix=True),
    dict(integrate_advection_by_parts=True),
    dict(beta=1.0),
    dict(integrate_advection_by_parts=True, beta=0.3),
    dict(have_absorption=True),
]


def variant_config(name, scale, variant):
    cfg = M.config(name, seed=7, scale=scale)
    cfg.source = np.cos(3 * cfg.mesh.node_pos[:, 0])
    cfg.absorption = 0.5 + cfg.mesh.node_pos[:, 1] if variant.get("have_absorption") else None
    return cfg


@pytest.mark.parametrize("name,scale", [("c1p2", 8), ("c3", 3)])
@pytest.mark.parametrize("variant", VARIANTS, ids=[",".join(v) or "plain" for v in VARIANTS])
def test_two_pass_equals_colour_and_oracle(require_gpu, name, scale, variant):
    cfg = variant_config(name, scale, variant)
    ref = O.reference_system_opts(cfg, variant)
    pr = cg.problem_from_config(cfg)
    out = {}
    for scheme in ("colour", "blocked"):
        for prec in ("fp64", "fp32"):
            print(f"\n[case] {name} scale {scale} {variant} {scheme} {prec}", flush=True)
            opts = cg.options_from_config(cfg, precision=prec, scheme=scheme)
            for k, v in variant.items():
                setattr(opts, k, v)
            cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
            cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
            out[(scheme, prec)] = pr.matrix()
    for prec in ("fp64", "fp32"):
        assert np.array_equal(out[("colour", prec)][0], out[("blocked", prec)][0]), prec
        assert np.array_equal(out[("colour", prec)][1], out[("blocked", prec)][1]), prec
    A, rhs, _ = out[("blocked", "fp64")]
    assert rel(A, ref["A"]) < FP64_RTOL
    assert rel(rhs, ref["rhs"]) < FP64_RTOL
    A32, rhs32, _ = out[("blocked", "fp32")]
    assert rel(A32, ref["A"]) < FP32_RTOL
    assert rel(rhs32, ref["rhs"]) < FP32_RTOL
    pr.close()


def test_two_pass_repeats_and_follows_field_changes(require_gpu):
    """The plan is built once; a second assembly with other field values
    gives the colour scheme's bits again, and the same call twice gives
    the same bits."""
    cfg = M.config("c3", seed=9, scale=3)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg, scheme="blocked")
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    first = pr.matrix()
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    again = pr.matrix()
    assert np.array_equal(first[0], again[0]) and np.array_equal(first[1], again[1])
    pr.set_fields(T=cfg.T * 2.0 + 1.0)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    two = pr.matrix()
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg, scheme="colour"), cfg.dt)
    col = pr.matrix()
    assert np.array_equal(two[0], col[0]) and np.array_equal(two[1], col[1])
    assert not np.array_equal(two[1], first[1])
    pr.close()


def test_two_pass_out_of_pattern_insertion_is_an_error(require_gpu):
    cfg = M.config("c1p2", seed=1, scale=4)
    sp = cg.make_sparsity(cfg.mesh)
    f = sp.findrm.copy()
    c = sp.colm.copy()
    drop = 1 if c[0] == 1 else 0  # an off-diagonal entry of row 1
    keep = np.ones(c.size, bool)
    keep[drop] = False
    f[1:] -= 1
    bad = cg.Sparsity(f, c[keep].copy(), np.where(sp.centrm > drop + 1, sp.centrm - 1, sp.centrm).astype(np.int32))
    pr = cg.CGProblem(cfg.mesh, quad_degree=cfg.quad_degree, sparsity=bad)
    pr.set_fields(T=cfg.T, velocity=cfg.U, diffusivity=1.0, source=cfg.source)
    opts = cg.options_from_config(cfg, scheme="blocked")
    print("\n[case] c1p2 two-pass, one entry dropped from the pattern", flush=True)
    with pytest.raises(FlgpuError, match="outside sparsity"):
        cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    pr.close()

"""GPU parity with an anisotropic, spatially varying diffusivity tensor
(dshape_tensor_dshape, femtools/FETools.F90:668-698, used by
add_diffusivity_element_cg, assemble/Advection_Diffusion_CG.F90:1195),
alone and under SUPG (whose nu_bar inverts the tensor), on every kernel
that takes it: the blocked scheme's gathered kernel, the P1 colour kernel,
the P2 entry kernel and the stabilised kernel.  1e-12 relative (FP64),
1e-5 (FP32)."""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, meshgen as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def tensor_field(pos, scale):
    """K(x) = R(x) diag(1, 0.3, 0.1) R(x)^T scale: symmetric positive
    definite, rotating with position."""
    n, d = pos.shape
    ang = 2.0 * pos[:, 0] + pos[:, 1]
    c, s = np.cos(ang), np.sin(ang)
    K = np.zeros((n, d, d))
    if d == 2:
        R = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], 1)
        D = np.array([1.0, 0.3])
    else:
        R = np.zeros((n, 3, 3))
        R[:, 0, 0], R[:, 0, 1], R[:, 1, 0], R[:, 1, 1], R[:, 2, 2] = c, -s, s, c, 1.0
        D = np.array([1.0, 0.3, 0.1])
    for k in range(n):
        K[k] = R[k] @ np.diag(D) @ R[k].T
    return K * scale


@pytest.mark.parametrize("name,scale,variant,precision", [
    ("c1", 12, {"scheme": "colour"}, "fp64"),
    ("c1", 12, {"scheme": "blocked"}, "fp64"),
    ("c2", 6, {"scheme": "auto"}, "fp64"),
    ("c2", 6, {"scheme": "colour"}, "fp32"),
    ("c3", 2, {"scheme": "colour"}, "fp64"),
    ("c1", 12, {"scheme": "auto", "stabilisation": "supg"}, "fp64"),
    ("c3", 2, {"scheme": "auto", "stabilisation": "supg", "nu_bar_scheme": "critical_rule"}, "fp64"),
])
def test_anisotropic_diffusivity_matches_oracle(require_gpu, name, scale, variant, precision):
    cfg = M.config(name, seed=21, scale=scale)
    cfg.kappa = tensor_field(cfg.mesh.node_pos, 1e-2 if cfg.U is not None else 1.0)
    var = dict(variant)
    scheme = var.pop("scheme")
    var["isotropic_diffusivity"] = False
    print(f"\n[case] {name} anisotropic {scheme} {var} {precision}", flush=True)
    ref = O.reference_system_opts(cfg, var)
    pr = cg.problem_from_config(cfg)
    opts = cg.options_from_config(cfg, precision=precision, scheme=scheme)
    for k, v in var.items():
        setattr(opts, k, v)
    cg.assemble_advection_diffusion_cg(pr, opts, cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, _ = pr.matrix()
    pr.close()
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert rel(A, ref["A"]) < tol
    assert rel(rhs, ref["rhs"]) < tol

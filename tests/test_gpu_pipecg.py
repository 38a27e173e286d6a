"""The single-reduction CG ("pipecg": the Chronopoulos & Gear recurrence)
against the (F)PCG loop and the oracle.

In exact arithmetic the recurrence is PCG: the same iterates, so the same
iteration count to a tolerance; in floating point the counts may differ by
a few.  For every preconditioner the GPU offers with CG (none, jacobi, sor,
fsai, learned), in FP64 and mixed precision: the solve reaches the FP64
true relative residual, its solution agrees with the oracle's FP64
Jacobi-PCG solution solved to 1e-13, and its iteration count is within
max(3, 5%) of the (F)PCG loop's in FP64.  In mixed precision the two FP32
recurrences drift apart differently between the FP64 residual
replacements (pipecg carries s = A p by recurrence, the (F)PCG loop forms
A p afresh), so the counts are held to max(6, 15%).
"""
import numpy as np
import pytest

import oracle as O
from fluidity_amd import cg, learned, meshgen as M

pytestmark = pytest.mark.gpu


def gpu_system(cfg):
    pr = cg.problem_from_config(cfg)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    return pr


CASES = [(name, scale, pc, prec) for name, scale in (("c2", 12), ("c4", 14))
         for pc in ("none", "jacobi", "sor", "fsai", "learned") for prec in ("fp64", "mixed")]


@pytest.mark.parametrize("name,scale,pc,prec", CASES, ids=[f"{c[0]}-{c[2]}-{c[3]}" for c in CASES])
def test_pipecg_matches_pcg(require_gpu, name, scale, pc, prec):
    cfg = M.config(name, seed=31, scale=scale)
    print(f"\n[case] pipecg {name} scale {scale} {pc} {prec}", flush=True)
    ref = O.reference_system(cfg)
    xe, _, reason_e, _, _ = O.petsc_solve("cg", "jacobi", 1e-13, ref["sparsity"], ref["A"], ref["inactive"],
                                          ref["rhs"], max_its=50000)
    assert reason_e > 0
    pr = gpu_system(cfg)
    if pc == "learned":
        learned.set_params(pr, learned.load())
    rtol = 1e-10
    base = "fcg" if pc == "learned" else "cg"
    x1, i1 = cg.petsc_solve(pr, cg.SolverOptions("pipecg", pc, relative_error=rtol, norm="unpreconditioned",
                                                 max_iterations=20000, precision=prec))
    _, i0 = cg.petsc_solve(pr, cg.SolverOptions(base, pc, relative_error=rtol, norm="unpreconditioned",
                                                max_iterations=20000, precision=prec))
    ex = np.max(np.abs(x1 - xe)) / np.max(np.abs(xe))
    print(f"[case] pipecg its {i1.its} reason {i1.reason} relres {i1.true_relres:.2e} x err {ex:.2e}; "
          f"{base} its {i0.its}", flush=True)
    assert i1.reason > 0 and i1.true_relres <= rtol * (1 + 1e-6)
    assert ex <= 1e-6
    slack = max(3, i0.its // 20) if prec == "fp64" else max(6, (3 * i0.its) // 20)
    assert abs(i1.its - i0.its) <= slack, (i1.its, i0.its)
    assert i1.fallback == 0
    pr.close()

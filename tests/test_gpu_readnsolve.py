"""petsc_readnsolve on the GPU (tools/petsc_readnsolve.F90:60-83 with the
device solver): a committed dump solves to scipy's solution, and a dump of
the 3-D system replays to the solution of the direct device solve."""
import os

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import oracle as O
from fluidity_amd import cg, meshgen as M
from fluidity_amd import readnsolve as R

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_8.matrixdump")


def test_fixture_solves_to_scipy(require_gpu):
    print("\n[case] readnsolve c1_8.matrixdump gmres", flush=True)
    f, c, v, b, _ = R.matrixdump_read(FIXTURE)
    n = f.size - 1
    A = sps.csr_matrix((v, c - 1, f - 1), shape=(n, n))
    want = spla.spsolve(A.tocsc(), b)
    solver = cg.SolverOptions(iterative_method="gmres", preconditioner="jacobi", relative_error=1e-12,
                              norm="unpreconditioned")
    x, info = R.petsc_readnsolve(FIXTURE, solver)
    assert info.reason > 0
    assert np.max(np.abs(x - want)) <= 1e-9 * np.max(np.abs(want))


def test_dump_replays_the_device_solve(require_gpu, tmp_path):
    print("\n[case] readnsolve c2 scale 8 cg", flush=True)
    cfg = M.config("c2", seed=2, scale=8)
    ref = O.reference_system(cfg)
    rowptr, col, val = O.reduced_system(ref["sparsity"], ref["A"], ref["inactive"])
    keep = val != 0.0
    row = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    lens = np.bincount(row[keep], minlength=rowptr.size - 1)
    findrm = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    path = str(tmp_path / "matrixdump")
    R.matrixdump_write(path, findrm, (col[keep] + 1).astype(np.int32), val[keep], ref["rhs"], None)
    solver = cg.SolverOptions(iterative_method="cg", preconditioner="jacobi", relative_error=1e-11,
                              norm="unpreconditioned")
    x_dump, info = R.petsc_readnsolve(path, solver)
    assert info.reason > 0
    pr = cg.problem_from_config(cfg)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    x_dir, info2 = cg.petsc_solve(pr, solver)
    pr.close()
    assert info2.reason > 0
    assert np.max(np.abs(x_dump - x_dir)) <= 1e-8 * np.max(np.abs(x_dir))


def test_failed_solve_dumps_the_petsc_system(require_gpu, tmp_path):
    """ConvergenceCheck: a solve stopped at max_its (reason -3) writes the
    system PETSc saw once; it is the oracle's reduced system, and solving
    it gives the converged device solve's solution off the Dirichlet rows."""
    print("\n[case] convergence_check dump c2 scale 6", flush=True)
    cfg = M.config("c2", seed=5, scale=6)
    ref = O.reference_system(cfg)
    pr = cg.problem_from_config(cfg)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    short = cg.SolverOptions(iterative_method="cg", preconditioner="jacobi", relative_error=1e-12, max_iterations=3,
                             check_every=1)
    x, info = cg.petsc_solve(pr, short)
    assert info.reason == -3
    path = str(tmp_path / "matrixdump")
    cg._matrixdumped = False
    assert cg.convergence_check(pr, info, "Tracer", filename=path) is True
    f, c, v, b, x0 = R.matrixdump_read(path)
    rowptr, col, val = O.reduced_system(ref["sparsity"], ref["A"], ref["inactive"])
    n = rowptr.size - 1
    dense_ref = sps.csr_matrix((val, col, rowptr), shape=(n, n)).toarray()
    dense = sps.csr_matrix((v, c - 1, f - 1), shape=(n, n)).toarray()
    assert np.max(np.abs(dense - dense_ref)) <= 1e-12 * np.max(np.abs(dense_ref))
    assert x0 is not None and not x0.any()
    good = cg.SolverOptions(iterative_method="cg", preconditioner="jacobi", relative_error=1e-11,
                            norm="unpreconditioned")
    xd, infod = R.petsc_readnsolve(path, good)
    xf, infof = cg.petsc_solve(pr, good)
    assert infod.reason > 0 and infof.reason > 0
    free = ref["inactive"] == 0
    assert np.max(np.abs(xd[free] - xf[free])) <= 1e-8 * np.max(np.abs(xf))
    # one dump per process, as the reference's saved flag
    os.remove(path)
    cg.convergence_check(pr, info, "Tracer", filename=path)
    assert not os.path.exists(path)
    pr.close()

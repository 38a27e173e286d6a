"""GPU tests of the learned preconditioner: forward factor and gradient
against the numpy restatement (oracle/learned_ref.py, BF16 operands
emulated), training, and its use inside flexible PCG with the Jacobi
fallback."""
import numpy as np
import pytest

import learned_ref as LR
import oracle as O
from fluidity_amd import cg, learned, meshgen as M

pytestmark = pytest.mark.gpu


def system(name="c4", scale=6, seed=3):
    cfg = M.config(name, seed=seed, scale=scale)
    pr = cg.problem_from_config(cfg)
    cg.assemble_advection_diffusion_cg(pr, cg.options_from_config(cfg), cfg.dt)
    cg.apply_dirichlet_conditions(pr, cfg.mesh.boundary, 0.0, cfg.dt)
    A, rhs, inact = pr.matrix()
    sp = (pr.sparsity.findrm, pr.sparsity.colm, pr.sparsity.centrm)
    g = LR.Graph(*O.reduced_system(sp, A, inact))
    return cfg, pr, g


def perturbed_params(seed=7):
    th = LR.init_params(seed).astype(np.float64)
    th[LR.WOA:LR.WOA + LR.H] += 0.05
    th[LR.WOB:LR.WOB + LR.H] -= 0.03
    th[LR.WD:LR.WD + LR.H] += 0.02
    th[LR.G1] = 0.3
    th[LR.BO] = -0.4
    return th.astype(np.float32)


def test_param_layout_and_init(require_gpu):
    cfg, pr, g = system(scale=4)
    assert learned.param_count() == LR.TOTAL
    learned.init_params(pr, 7)
    assert np.array_equal(learned.get_params(pr), LR.init_params(7))
    pr.close()


def test_forward_factor_matches_oracle(require_gpu):
    cfg, pr, g = system(scale=6)
    th = perturbed_params()
    learned.set_params(pr, th)
    gv = learned.factor(pr)
    cache = LR.forward(th, g, rb=True)
    gref, _, _ = LR.factor(th, g, cache)
    err = np.max(np.abs(gv - gref)) / np.max(np.abs(gref))
    print("factor rel err", err)
    # FP32 vs FP64 sums can move an activation across a bfloat16 rounding
    # boundary (one 2^-8 step of one MFMA operand), hence 5e-3 here while
    # the loss and gradients agree to ~1e-5
    assert err < 5e-3
    pr.close()


def test_pc_apply_matches_exported_factor(require_gpu):
    """flgpu_pc_apply with the learned preconditioner (the solver's FP64
    halves, the upper one gathered from the transposed lower values) equals
    G^T (G r) formed with the exported factor in A's pattern (FP32
    values, hence 1e-5)."""
    cfg, pr, g = system(scale=8)
    learned.set_params(pr, perturbed_params())
    gv = learned.factor(pr).astype(np.float64)
    f, c = pr.sparsity.findrm - 1, pr.sparsity.colm - 1
    n = f.size - 1
    row = np.repeat(np.arange(n), np.diff(f))
    low = c <= row
    r = np.random.default_rng(5).standard_normal(n)
    y = np.zeros(n)
    np.add.at(y, row[low], gv[low] * r[c[low]])
    z_ref = np.zeros(n)
    np.add.at(z_ref, c[low], gv[low] * y[row[low]])
    z = np.zeros(n)
    import ctypes as C
    from fluidity_amd import lib as L
    L.check(pr.lib.flgpu_pc_apply(pr.ctx, C.byref(cg.SolverOptions("cg", "learned").params()), L.ptr(r), L.ptr(z)),
            "flgpu_pc_apply")
    err = np.max(np.abs(z - z_ref)) / np.max(np.abs(z_ref))
    print("learned pc_apply rel err", err)
    assert err < 1e-5
    pr.close()


def test_gradient_matches_oracle(require_gpu):
    cfg, pr, g = system(scale=6)
    th = perturbed_params()
    learned.set_params(pr, th)
    loss, grad = learned.loss_grad(pr, nprobe=8, seed=42)
    z = LR.probes(g.n, 8, 42)
    lref, gref, _ = LR.loss_and_grad(th, g, z, rb=True)
    print("loss", loss, lref)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    scale = np.max(np.abs(gref))
    err = np.max(np.abs(grad - gref)) / scale
    # per-block check so that a wrong layer does not hide behind a large one
    for lo, hi in [(0, LR.L0), (LR.layer(0), LR.layer(1)), (LR.layer(1), LR.layer(2)),
                   (LR.layer(2), LR.OUT), (LR.OUT, LR.TOTAL)]:
        s = max(np.max(np.abs(gref[lo:hi])), 1e-30)
        e = np.max(np.abs(grad[lo:hi] - gref[lo:hi])) / s
        print("block", lo, hi, "rel err", e)
        assert e < 5e-3
    assert err < 2e-3
    pr.close()


def test_training_reduces_loss_and_iterations(require_gpu):
    cfg, pr, g = system("c4", scale=10, seed=5)
    learned.init_params(pr, 1)
    so_j = cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned")
    _, ij = cg.petsc_solve(pr, so_j)
    losses = [learned.train_step(pr, lr=3e-3, nprobe=8, seed=s) for s in range(120)]
    print("loss first/last", losses[0], np.mean(losses[-10:]))
    assert np.mean(losses[-10:]) < 0.8 * losses[0]
    so_l = cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned")
    x, il = cg.petsc_solve(pr, so_l)
    print("jacobi its", ij.its, "learned its", il.its, "fallback", il.fallback)
    assert il.reason > 0 and il.true_relres <= 1e-8
    assert il.fallback == 0
    assert il.its < ij.its
    so_m = cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned", precision="mixed")
    xm, im = cg.petsc_solve(pr, so_m)
    assert im.reason > 0 and im.true_relres <= 1e-8
    assert np.max(np.abs(xm - x)) <= 1e-6 * np.max(np.abs(x))
    pr.close()


def test_stalled_network_falls_back_to_jacobi(require_gpu):
    """A network whose factor is far too small stalls the flexible PCG; the
    solver switches to Jacobi and still reaches the tolerance."""
    cfg, pr, g = system("c4", scale=8, seed=6)
    th = np.zeros(LR.TOTAL, np.float32)
    th[LR.BD] = -12.0            # Gs_ii = exp(-12): the preconditioner is ~0
    th[LR.BO] = 40.0             # and wildly wrong off the diagonal
    learned.set_params(pr, th)
    so = cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned",
                          stall_window=32, stall_factor=0.9, check_every=16)
    x, info = cg.petsc_solve(pr, so)
    print("its", info.its, "fallback", info.fallback, "reason", info.reason, "relres", info.true_relres)
    assert info.fallback == 1
    assert info.reason > 0 and info.true_relres <= 1e-8
    pr.close()


@pytest.mark.parametrize("name,scale,seed", [("c4", 12, 9101), ("c5", 14, 9102)])
def test_committed_weights_on_unseen_system(require_gpu, name, scale, seed):
    """The shipped weights (trained on the GPU by fluidity_amd.train_learned)
    on a system none of the training seeds made: learned FCG converges to
    the true residual without falling back, in fewer iterations than
    Jacobi-PCG, and agrees with it."""
    cfg, pr, g = system(name, scale=scale, seed=seed)
    learned.set_params(pr, learned.load())
    print(f"\n[case] committed weights on {name} scale {scale} seed {seed}", flush=True)
    so_j = cg.SolverOptions("cg", "jacobi", relative_error=1e-8, norm="unpreconditioned")
    xj, ij = cg.petsc_solve(pr, so_j)
    so_l = cg.SolverOptions("fcg", "learned", relative_error=1e-8, norm="unpreconditioned")
    xl, il = cg.petsc_solve(pr, so_l)
    print("jacobi its", ij.its, "learned its", il.its, "fallback", il.fallback)
    assert il.reason > 0 and il.true_relres <= 1e-8
    assert il.fallback == 0
    assert il.its < ij.its
    assert np.max(np.abs(xl - xj)) <= 1e-6 * np.max(np.abs(xj))
    pr.close()

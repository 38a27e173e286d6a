"""CPU checks of the network restatement used as the gradient oracle:
its analytic gradient against central finite differences of its own loss
(smooth FP64 mode), and the BF16 rounding helper."""
import numpy as np
import pytest

import learned_ref as LR
import oracle as O
from fluidity_amd import meshgen as M


@pytest.fixture(scope="module")
def small_graph():
    cfg = M.config("c4", seed=2, scale=3)
    ref = O.reference_system(cfg)
    rowptr, col, val = O.reduced_system(ref["sparsity"], ref["A"], ref["inactive"])
    return LR.Graph(rowptr, col, val)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.1415926], dtype=np.float32)
    r = LR.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == np.float32(1.0 + 2.0 ** -7)
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-7


def test_transpose_map_is_an_involution(small_graph):
    g = small_graph
    assert np.array_equal(g.tmap[g.tmap], np.arange(g.col.size))
    assert np.array_equal(g.row[g.tmap], g.col)


def test_analytic_gradient_matches_finite_differences(small_graph):
    g = small_graph
    th = LR.init_params(7).astype(np.float64)
    # move the output layer off zero so every parameter has a gradient
    th[LR.WOA:LR.WOA + LR.H] += 0.05
    th[LR.WOB:LR.WOB + LR.H] -= 0.03
    th[LR.WD:LR.WD + LR.H] += 0.02
    th[LR.G1] = 0.3
    th[LR.BO] = -0.4
    z = LR.probes(g.n, 8, 99)
    loss, grad, _ = LR.loss_and_grad(th, g, z, rb=False)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(LR.TOTAL, 40, replace=False),
                          [LR.G1, LR.G2, LR.BO, LR.BD, LR.WOA, LR.W_IN, LR.layer(2) + LR.C1]])
    for k in idx:
        e = 1e-6 * max(1.0, abs(th[k]))
        tp, tm = th.copy(), th.copy()
        tp[k] += e
        tm[k] -= e
        lp, _, _ = LR.loss_and_grad(tp, g, z, rb=False)
        lm, _, _ = LR.loss_and_grad(tm, g, z, rb=False)
        fd = (lp - lm) / (2 * e)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, np.max(np.abs(grad))) + 1e-4 * abs(fd), (k, fd, grad[k])


def test_zero_network_is_jacobi(small_graph):
    """With every output weight zero the factor is D^-1/2 (Jacobi), so the
    loss equals ||(I - D^-1 A) z||^2 / (N P)."""
    g = small_graph
    th = np.zeros(LR.TOTAL)
    z = LR.probes(g.n, 4, 5)
    loss, _, gv = LR.loss_and_grad(th, g, z, rb=False)
    import scipy.sparse as sps
    A = sps.csr_matrix((g.val, g.col, g.rowptr), shape=(g.n, g.n))
    d = A.diagonal()
    ref = np.sum((z - (A @ z) / d[:, None]) ** 2) / z.size
    assert abs(loss - ref) <= 1e-12 * ref

"""TEST INFRASTRUCTURE ONLY: numpy restatement of the learned
preconditioner's network, loss and analytic gradient (the checker of the
gfx950 forward/backward kernels in fluidity_amd/csrc/learned.hip).

bf16=True rounds exactly the operands the MFMA kernels round to bfloat16
(node-product inputs and weights, weight-gradient inputs), so the GPU and
this restatement differ only by FP32 summation order.  bf16=False gives
the smooth FP64 function used to pin the analytic gradient with finite
differences.
"""
from __future__ import annotations

import numpy as np

H, NFEAT, NLAYER = 32, 4, 3

# parameter layout (same offsets as the device code)
W_IN = 0
B_IN = W_IN + H * NFEAT
L0 = B_IN + H
WA, WB = 0, H * H
C1 = WB + H * H
C2 = C1 + H
BM = C2 + H
WU = BM + H
BU = WU + H * 2 * H
LSIZE = BU + H
OUT = L0 + NLAYER * LSIZE
WOA = OUT
WOB = WOA + H
WD = WOB + H
G1 = WD + H
G2 = G1 + 1
BO = G2 + 1
BD = BO + 1
TOTAL = BD + 1


def layer(l):
    return L0 + l * LSIZE


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rounding) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def smix(x):
    with np.errstate(over="ignore"):
        x = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def probes(n, P, seed):
    t = np.arange(n * P, dtype=np.uint64)
    z = np.where((smix(np.uint64(seed) ^ t) >> np.uint64(63)) == 1, 1.0, -1.0)
    return z.reshape(n, P)


class Graph:
    """Reduced system in CSR (0-based) with the derived scaled entries."""

    def __init__(self, rowptr, col, val):
        self.rowptr, self.col, self.val = rowptr, col, val
        self.n = rowptr.size - 1
        self.row = np.repeat(np.arange(self.n), np.diff(rowptr))
        dpos = np.full(self.n, -1)
        isd = self.col == self.row
        dpos[self.row[isd]] = np.nonzero(isd)[0]
        self.diag = dpos
        a = np.where(dpos >= 0, val[np.maximum(dpos, 0)], 1.0)
        self.dsq = np.where(a > 0, 1.0 / np.sqrt(np.where(a > 0, a, 1.0)), 1.0)
        self.s = val * self.dsq[self.row] * self.dsq[self.col]
        off = self.col != self.row
        self.off = off & (self.s != 0)
        rowsum = np.bincount(self.row, weights=self.s, minlength=self.n)
        nb = np.bincount(self.row[self.off], minlength=self.n)
        maxoff = np.zeros(self.n)
        np.maximum.at(maxoff, self.row[self.off], np.abs(self.s[self.off]))
        self.feat = np.stack([rowsum, nb / 16.0, maxoff, (nb == 0).astype(float)], axis=1)
        # transposed positions
        key = self.row * self.n + self.col
        tkey = self.col * self.n + self.row
        order = np.argsort(key)
        pos = np.searchsorted(key[order], tkey)
        self.tmap = order[np.minimum(pos, key.size - 1)]
        self.lower = self.col < self.row
        self.edge = self.col != self.row  # message edges (i <- j), j != i


def _gemm(X, W, rb):
    """X (N x K) times W^T (W: Dout x K), with bf16 operands when rb."""
    if rb:
        return bf16(X).astype(np.float64) @ bf16(W).astype(np.float64).T
    return X @ W.T


def forward(th, g: Graph, rb=True):
    th = np.asarray(th, dtype=np.float64)
    cache = {"h": [], "PQ": [], "M": [], "U": []}
    h = g.feat @ th[W_IN:W_IN + H * NFEAT].reshape(H, NFEAT).T + th[B_IN:B_IN + H]
    cache["h"].append(h)
    e = g.edge
    ri, cj, s = g.row[e], g.col[e], g.s[e]
    for l in range(NLAYER):
        lp = th[layer(l):layer(l) + LSIZE]
        Wab = np.concatenate([lp[WA:WA + H * H].reshape(H, H), lp[WB:WB + H * H].reshape(H, H)], axis=0)
        PQ = _gemm(h, Wab, rb)
        pre = PQ[ri, :H] + PQ[cj, H:] + np.outer(s, lp[C1:C1 + H]) + np.outer(np.abs(s), lp[C2:C2 + H]) + lp[BM:BM + H]
        M = np.zeros((g.n, H))
        np.add.at(M, ri, np.maximum(pre, 0.0))
        Wu = lp[WU:WU + H * 2 * H].reshape(H, 2 * H)
        U = _gemm(np.concatenate([h, M], axis=1), Wu, rb) + lp[BU:BU + H]
        h = h + np.maximum(U, 0.0)
        cache["PQ"].append(PQ)
        cache["M"].append(M)
        cache["U"].append(U)
        cache["h"].append(h)
    alpha = h @ th[WOA:WOA + H]
    beta = h @ th[WOB:WOB + H]
    psi = h @ th[WD:WD + H] + th[BD]
    cache.update(alpha=alpha, beta=beta, psi=psi)
    return cache


def factor(th, g: Graph, cache):
    """Values of the preconditioner in A's pattern (lower: Gs_ij d_j,
    upper: d_i Gs_ji, diagonal: Gs_ii d_i) and Gs on lower+diag."""
    th = np.asarray(th, dtype=np.float64)
    a, b, psi = cache["alpha"], cache["beta"], cache["psi"]
    s = g.s
    phi = a[g.row] + b[g.col] + th[G1] * s + th[G2] * np.abs(s) + th[BO]
    gs = np.where(g.lower, s * phi, 0.0)
    gs = np.where(g.col == g.row, np.exp(psi[g.row]), gs)
    gv = np.where(g.col <= g.row, gs * g.dsq[g.col], 0.0)
    upper = g.col > g.row
    gv = np.where(upper, g.dsq[g.row] * gs[g.tmap], gv)
    return gv, gs, phi


def _spmv(g: Graph, vals, x, part):
    mask = {0: np.ones_like(g.lower), 1: g.col <= g.row, 2: g.col >= g.row}[part]
    y = np.zeros((g.n, x.shape[1]))
    np.add.at(y, g.row[mask], vals[mask][:, None] * x[g.col[mask]])
    return y


def loss_and_grad(th, g: Graph, z, rb=True):
    th = np.asarray(th, dtype=np.float64)
    cache = forward(th, g, rb)
    gv, gs, phi = factor(th, g, cache)
    n, P = z.shape
    u = _spmv(g, g.val, z, 0)
    v = _spmv(g, gv, u, 1)
    w = _spmv(g, gv, v, 2)
    r = z - w
    scale = 1.0 / (n * P)
    loss = float(np.sum(r * r) * scale)
    dw = -2.0 * r * scale
    dv = _spmv(g, gv, dw, 1)
    grad = np.zeros(TOTAL)
    # d Gs on lower + diagonal entries
    lowd = g.col <= g.row
    dgs = np.where(lowd, np.sum(dv[g.row] * u[g.col] + dw[g.col] * v[g.row], axis=1) * g.dsq[g.col], 0.0)
    dphi = np.where(g.lower, dgs * g.s, 0.0)
    isd = g.col == g.row
    dpsi = np.zeros(n)
    dpsi[g.row[isd]] = dgs[isd] * gs[isd]
    dalpha = np.bincount(g.row, weights=dphi, minlength=n)
    dbeta = np.bincount(g.col, weights=dphi, minlength=n)
    grad[G1] = np.sum(dphi * g.s)
    grad[G2] = np.sum(dphi * np.abs(g.s))
    grad[BO] = np.sum(dphi)
    hL = cache["h"][-1]
    grad[WOA:WOA + H] = dalpha @ hL
    grad[WOB:WOB + H] = dbeta @ hL
    grad[WD:WD + H] = dpsi @ hL
    grad[BD] = np.sum(dpsi)
    dh = np.outer(dalpha, th[WOA:WOA + H]) + np.outer(dbeta, th[WOB:WOB + H]) + np.outer(dpsi, th[WD:WD + H])
    e = g.edge
    ri, cj, s = g.row[e], g.col[e], g.s[e]
    for l in reversed(range(NLAYER)):
        lp = th[layer(l):layer(l) + LSIZE]
        o = layer(l)
        hin, PQ, M, U = cache["h"][l], cache["PQ"][l], cache["M"][l], cache["U"][l]
        dU = dh * (U > 0)
        X = np.concatenate([hin, M], axis=1)
        grad[o + WU:o + WU + H * 2 * H] += (_gemm(dU.T, X.T, rb)).ravel()  # dU^T X
        grad[o + BU:o + BU + H] += dU.sum(axis=0)
        Wu = lp[WU:WU + H * 2 * H].reshape(H, 2 * H)
        dHM = _gemm(dU, Wu.T, rb)
        dh = dh + dHM[:, :H]
        dM = dHM[:, H:]
        pre = PQ[ri, :H] + PQ[cj, H:] + np.outer(s, lp[C1:C1 + H]) + np.outer(np.abs(s), lp[C2:C2 + H]) + lp[BM:BM + H]
        dr = dM[ri] * (pre > 0)
        dP = np.zeros((n, H))
        dQ = np.zeros((n, H))
        np.add.at(dP, ri, dr)
        np.add.at(dQ, cj, dr)
        grad[o + C1:o + C1 + H] += (dr * s[:, None]).sum(axis=0)
        grad[o + C2:o + C2 + H] += (dr * np.abs(s)[:, None]).sum(axis=0)
        grad[o + BM:o + BM + H] += dr.sum(axis=0)
        dPQ = np.concatenate([dP, dQ], axis=1)
        grad[o + WA:o + WA + 2 * H * H] += _gemm(dPQ.T, hin.T, rb).ravel()
        Wab = np.concatenate([lp[WA:WA + H * H].reshape(H, H), lp[WB:WB + H * H].reshape(H, H)], axis=0)
        dh = dh + _gemm(dPQ, Wab.T, rb)
    grad[W_IN:W_IN + H * NFEAT] = (dh.T @ g.feat).ravel()
    grad[B_IN:B_IN + H] = dh.sum(axis=0)
    return loss, grad, gv


def init_params(seed):
    """Same deterministic initialisation as the device library."""
    th = np.zeros(TOTAL, dtype=np.float32)
    q = [1]

    def rnd():
        with np.errstate(over="ignore"):
            x = np.uint64(seed) ^ (np.uint64(q[0]) * np.uint64(0x9E3779B97F4A7C15))
        q[0] += 1
        v = smix(np.array([x], dtype=np.uint64))[0]
        return float((v >> np.uint64(11))) * (1.0 / 9007199254740992.0) * 2.0 - 1.0

    for i in range(H * NFEAT):
        th[W_IN + i] = rnd() * np.sqrt(3.0 / NFEAT)
    for l in range(NLAYER):
        o = layer(l)
        for i in range(H * H):
            th[o + WA + i] = rnd() * np.sqrt(3.0 / H)
        for i in range(H * H):
            th[o + WB + i] = rnd() * np.sqrt(3.0 / H)
        for i in range(H):
            th[o + C1 + i] = rnd()
        for i in range(H):
            th[o + C2 + i] = rnd()
        for i in range(H * 2 * H):
            th[o + WU + i] = rnd() * np.sqrt(3.0 / (2 * H)) * 0.5
    for i in range(H):
        th[WOA + i] = rnd() * 0.01
    for i in range(H):
        th[WOB + i] = rnd() * 0.01
    for i in range(H):
        th[WD + i] = rnd() * 0.01
    return th

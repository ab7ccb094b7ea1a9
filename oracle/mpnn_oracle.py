"""CPU restatement of the learned preconditioner -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker of the device library (mpi-cst-fem-solver_amd/csrc/mpnn.hip).

The reference solver has no learned preconditioner; this module restates the device
network so its outputs and gradients can be checked:

  features (9 live of 16): 1, std and range of ln E over the node's elements, the two
      Dirichlet flags, a01/sqrt(a00 a11), ln(a00/a11), degree/6,
      ln tr K_ii - mean_{j ~ i} ln tr K_jj
  h0 = relu(f W0 + b0); 3 rounds h = relu([h, mean_{j ~ i, j != i} h_j] Wl + bl)
  decoder per stored block q = (i, j): [a, b, d, c] = W^T (h_i * h_j) + bias
      (diagonal / off-diagonal weights)
  M_ij = S_i [[a, b], [b, d]] S_j + c D_i^-1 K_ij D_j^-1, D = diag(K), S = D^-1/2
  loss = (1/K) sum_k |M r_k - e_k|_A^2 / (e_k' r_k)

Two precisions: "f64" (everything FP64, no rounding -- the analytic backward is pinned
against central finite differences in this mode) and "bf16" (the device's rounding points:
features, weights, activations and the ReLU-masked gradients rounded to BF16, products
accumulated in FP32), used to compare with the device within BF16 tolerances.

Pinning (tests/test_oracle.py, tests/test_gpu_mpnn.py):
  - backward vs central finite differences of the loss (f64 mode), every parameter block;
  - the initial decoder (W = 0, diagonal bias (1, 0, 1, 0)) gives M = diag(K)^-1 exactly;
  - decoder biases (alpha, 0, alpha, -beta) give the polynomial D^-1 (alpha I - beta K D^-1);
  - M is symmetric for any parameters;
  - the loss is 0 when M r_k = e_k (pairs made from M's own inverse action).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

H = 64
F = 16
LIVE = 9
ROUNDS = 3
DEC = 4     # decoder outputs per block: a, b, d, c

W0 = 0
B0 = W0 + F * H
WL = B0 + H
BL = WL + ROUNDS * 2 * H * H
WD = BL + ROUNDS * H
BD = WD + H * DEC
WO = BD + DEC
BO = WO + H * DEC
NPARAMS = BO + DEC


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    """Decoder at the Jacobi state; hidden layers random (any values are valid here)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    th = np.zeros(NPARAMS, dtype=np.float32)
    w0 = np.zeros((F, H), dtype=np.float32)
    w0[:LIVE] = rng.standard_normal((LIVE, H)) * np.sqrt(2.0 / LIVE)
    th[W0:B0] = w0.reshape(-1)
    for l in range(ROUNDS):
        o = WL + l * 2 * H * H
        th[o:o + 2 * H * H] = rng.standard_normal(2 * H * H) * np.sqrt(2.0 / (2 * H))
    th[BD] = 1.0
    th[BD + 2] = 1.0
    return th


class Graph:
    """Block structure of a mesh system: node adjacency from the scalar CSR pattern (rows 2i
    hold columns 2j, 2j+1 of every adjacent node j, ascending), 2x2 value blocks, node ->
    elements, per-element modulus and the Dirichlet flags."""

    def __init__(self, row_ptr, cols, vals, tri, nn, E_elem=None, fixed=None):
        row_ptr = np.asarray(row_ptr, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        self.nb = nn
        self.n = 2 * nn
        deg = (row_ptr[2 * np.arange(nn) + 1] - row_ptr[2 * np.arange(nn)]) // 2
        self.brow_ptr = np.zeros(nn + 1, dtype=np.int64)
        self.brow_ptr[1:] = np.cumsum(deg)
        nblk = int(self.brow_ptr[-1])
        self.brow = np.repeat(np.arange(nn), deg)
        self.bcol = np.empty(nblk, dtype=np.int64)
        blocks = np.empty((nblk, 2, 2))
        for i in range(nn):
            a0, a1 = row_ptr[2 * i], row_ptr[2 * i + 1]
            b0 = row_ptr[2 * i + 1]
            q0, q1 = self.brow_ptr[i], self.brow_ptr[i + 1]
            c = cols[a0:a1]
            self.bcol[q0:q1] = c[0::2] // 2
            blocks[q0:q1, 0, 0] = vals[a0:a1][0::2]
            blocks[q0:q1, 0, 1] = vals[a0:a1][1::2]
            blocks[q0:q1, 1, 0] = vals[b0:b0 + (a1 - a0)][0::2]
            blocks[q0:q1, 1, 1] = vals[b0:b0 + (a1 - a0)][1::2]
        self.blocks = blocks
        self.diag_q = np.array([self.brow_ptr[i] + np.searchsorted(
            self.bcol[self.brow_ptr[i]:self.brow_ptr[i + 1]], i) for i in range(nn)])
        # transpose position of every block
        key = self.brow * nn + self.bcol
        tkey = self.bcol * nn + self.brow
        order = np.argsort(key)
        self.btrans = order[np.searchsorted(key[order], tkey)]
        self.offdiag = self.brow != self.bcol
        self.nbr_count = np.bincount(self.brow[self.offdiag], minlength=nn)
        self.A = sp.csr_matrix((vals, cols, row_ptr), shape=(self.n, self.n))
        d = np.concatenate([blocks[self.diag_q, 0, 0][:, None], blocks[self.diag_q, 1, 1][:, None]],
                           axis=1).reshape(-1)
        self.invd = 1.0 / d
        tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
        self.tri = tri
        self.E_elem = None if E_elem is None else np.asarray(E_elem, dtype=np.float64)
        self.fixed = np.zeros(self.n) if fixed is None else np.asarray(fixed, dtype=np.float64)

    def features(self) -> np.ndarray:
        nn = self.nb
        f = np.zeros((nn, F))
        f[:, 0] = 1.0
        if self.E_elem is not None:
            lnE = np.log(self.E_elem)
            node = self.tri.reshape(-1)
            val = np.repeat(lnE, 3)
            cnt = np.bincount(node, minlength=nn)
            s = np.bincount(node, weights=val, minlength=nn)
            s2 = np.bincount(node, weights=val * val, minlength=nn)
            lo = np.full(nn, np.inf)
            hi = np.full(nn, -np.inf)
            np.minimum.at(lo, node, val)
            np.maximum.at(hi, node, val)
            has = cnt > 0
            mean = np.where(has, s / np.maximum(cnt, 1), 0.0)
            f[:, 1] = np.where(has, np.sqrt(np.maximum(s2 / np.maximum(cnt, 1) - mean * mean, 0.0)), 0.0)
            f[:, 2] = np.where(has, hi - lo, 0.0)
        f[:, 3] = self.fixed[0::2]
        f[:, 4] = self.fixed[1::2]
        dg = self.blocks[self.diag_q]
        a00, a01, a11 = dg[:, 0, 0], dg[:, 0, 1], dg[:, 1, 1]
        f[:, 5] = a01 / np.sqrt(a00 * a11)
        f[:, 6] = np.log(a00 / a11)
        f[:, 7] = (self.brow_ptr[1:] - self.brow_ptr[:-1] - 1) / 6.0
        lt = np.log(a00 + a11)
        od = self.offdiag
        s = np.bincount(self.brow[od], weights=lt[self.bcol[od]], minlength=nn)
        cnt = self.nbr_count
        f[:, 8] = np.where(cnt > 0, lt - s / np.maximum(cnt, 1), 0.0)
        return f

    def aggregate(self, h, dtype):
        od = self.offdiag
        s = np.zeros(h.shape, dtype=dtype)
        np.add.at(s, self.brow[od], h[self.bcol[od]].astype(dtype))
        cnt = np.maximum(self.nbr_count, 1).astype(dtype)
        return np.where(self.nbr_count[:, None] > 0, s / cnt[:, None], 0.0).astype(dtype)

    def aggregate_bwd(self, dm):
        """Adjoint of aggregate: dh_j = sum_{i ~ j, i != j} dm_i / deg_i."""
        od = self.offdiag
        out = np.zeros(dm.shape)
        w = dm[self.brow[od]] / np.maximum(self.nbr_count[self.brow[od]], 1)[:, None]
        np.add.at(out, self.bcol[od], w)
        return out


def _relu(x):
    return np.maximum(x, 0.0)


def forward(theta, g: Graph, prec="f64"):
    """Returns (ghat [nblk x 3], cache)."""
    lowp = prec == "bf16"
    rnd = bf16 if lowp else (lambda a: np.asarray(a, dtype=np.float64))
    dt = np.float32 if lowp else np.float64
    th = np.asarray(theta, dtype=np.float32 if lowp else np.float64)
    feat = rnd(g.features().astype(np.float32) if lowp else g.features())
    W0m = rnd(th[W0:B0].reshape(F, H))
    Hs = [rnd(_relu(feat.astype(dt) @ W0m.astype(dt) + th[B0:WL].astype(dt)))]
    Xs = [None]
    for l in range(ROUNDS):
        h = Hs[-1]
        m = rnd(g.aggregate(h.astype(dt), dt))
        X = np.concatenate([h, m], axis=1)
        Wl = rnd(th[WL + l * 2 * H * H:WL + (l + 1) * 2 * H * H].reshape(2 * H, H))
        bl = th[BL + l * H:BL + (l + 1) * H].astype(dt)
        Xs.append(X)
        Hs.append(rnd(_relu(X.astype(dt) @ Wl.astype(dt) + bl)))
    h = Hs[-1].astype(dt)
    p = h[g.brow] * h[g.bcol]
    Wd = th[WD:BD].reshape(H, DEC).astype(dt)
    Wo = th[WO:BO].reshape(H, DEC).astype(dt)
    diag = ~g.offdiag
    ghat = np.where(diag[:, None], p @ Wd + th[BD:WO].astype(dt), p @ Wo + th[BO:].astype(dt))
    return ghat, dict(feat=feat, H=Hs, X=Xs, p=p, th=th, prec=prec)


def scaled_blocks(g: Graph):
    """D_i^-1 K_ij D_j^-1 per stored block (FP64)."""
    di = g.invd.reshape(-1, 2)
    return di[g.brow][:, :, None] * g.blocks * di[g.bcol][:, None, :]


def m_blocks(ghat, g: Graph):
    """2x2 blocks of M (FP64)."""
    s = np.sqrt(g.invd)
    si0, si1 = s[2 * g.brow], s[2 * g.brow + 1]
    sj0, sj1 = s[2 * g.bcol], s[2 * g.bcol + 1]
    a, b, d, c = (ghat[:, u].astype(np.float64) for u in range(DEC))
    mb = np.empty((len(a), 2, 2))
    mb[:, 0, 0] = si0 * a * sj0
    mb[:, 0, 1] = si0 * b * sj1
    mb[:, 1, 0] = si1 * b * sj0
    mb[:, 1, 1] = si1 * d * sj1
    return mb + c[:, None, None] * scaled_blocks(g)


def m_matrix(ghat, g: Graph):
    mb = m_blocks(ghat, g)
    rows = np.concatenate([2 * g.brow, 2 * g.brow, 2 * g.brow + 1, 2 * g.brow + 1])
    cols = np.concatenate([2 * g.bcol, 2 * g.bcol + 1, 2 * g.bcol, 2 * g.bcol + 1])
    v = np.concatenate([mb[:, 0, 0], mb[:, 0, 1], mb[:, 1, 0], mb[:, 1, 1]])
    return sp.csr_matrix((v, (rows, cols)), shape=(g.n, g.n))


def loss_and_grad(theta, g: Graph, R, E, prec="f64"):
    """Loss over the pairs (rows of R, E) and its gradient (FP64)."""
    ghat, cache = forward(theta, g, prec)
    M = m_matrix(ghat, g)
    R = np.asarray(R, dtype=np.float64)
    E = np.asarray(E, dtype=np.float64)
    K = R.shape[0]
    loss = 0.0
    V = np.empty_like(R)
    w = np.empty(K)
    for k in range(K):
        d = M @ R[k] - E[k]
        v = g.A @ d
        eae = E[k] @ R[k]
        ce = 1.0 / eae if eae > 0 else 0.0
        loss += (d @ v) * ce / K
        V[k] = v
        w[k] = 2.0 * ce / K
    # dL/dM_ij = sum_k w_k v_k[i] r_k[j]^T, mapped onto [a, b, d] through M = S G S
    s = np.sqrt(g.invd)
    i, j = g.brow, g.bcol
    Vw = V * w[:, None]
    g00 = np.einsum("ki,ki->i", Vw[:, 2 * i], R[:, 2 * j])
    g01 = np.einsum("ki,ki->i", Vw[:, 2 * i], R[:, 2 * j + 1])
    g10 = np.einsum("ki,ki->i", Vw[:, 2 * i + 1], R[:, 2 * j])
    g11 = np.einsum("ki,ki->i", Vw[:, 2 * i + 1], R[:, 2 * j + 1])
    sb = scaled_blocks(g)
    dghat = np.stack([g00 * s[2 * i] * s[2 * j],
                      g01 * s[2 * i] * s[2 * j + 1] + g10 * s[2 * i + 1] * s[2 * j],
                      g11 * s[2 * i + 1] * s[2 * j + 1],
                      g00 * sb[:, 0, 0] + g01 * sb[:, 0, 1] + g10 * sb[:, 1, 0]
                      + g11 * sb[:, 1, 1]], axis=1)
    grad = backward(dghat, g, cache)
    return loss, grad, dict(ghat=ghat, dghat=dghat)


def backward(dghat, g: Graph, cache):
    th = np.asarray(cache["th"], dtype=np.float64)
    lowp = cache["prec"] == "bf16"
    rnd = bf16 if lowp else (lambda a: a)
    grad = np.zeros(NPARAMS)
    Hs, Xs = cache["H"], cache["X"]
    h = Hs[-1].astype(np.float64)
    p = h[g.brow] * h[g.bcol]
    diag = ~g.offdiag
    grad[WD:BD] = (p[diag].T @ dghat[diag]).reshape(-1)
    grad[BD:WO] = dghat[diag].sum(axis=0)
    grad[WO:BO] = (p[g.offdiag].T @ dghat[g.offdiag]).reshape(-1)
    grad[BO:] = dghat[g.offdiag].sum(axis=0)
    Wd = th[WD:BD].reshape(H, DEC)
    Wo = th[WO:BO].reshape(H, DEC)
    Wq = np.where(diag[:, None, None], Wd[None], Wo[None])     # nblk x H x DEC
    gq = np.einsum("qco,qo->qc", Wq, dghat)                    # W dghat_q
    dH = np.zeros_like(h)
    np.add.at(dH, g.brow, gq * h[g.bcol])
    np.add.at(dH, g.bcol, gq * h[g.brow])
    for l in range(ROUNDS, 0, -1):
        dZ = rnd(np.where(Hs[l] > 0, dH, 0.0)).astype(np.float64)
        X = Xs[l].astype(np.float64)
        o = WL + (l - 1) * 2 * H * H
        grad[o:o + 2 * H * H] = (X.T @ dZ).reshape(-1)
        grad[BL + (l - 1) * H:BL + l * H] = dZ.sum(axis=0)
        Wl = rnd(th[o:o + 2 * H * H].reshape(2 * H, H)).astype(np.float64)
        dX = dZ @ Wl.T
        dH = dX[:, :H] + g.aggregate_bwd(dX[:, H:])
    dZ = rnd(np.where(Hs[0] > 0, dH, 0.0)).astype(np.float64)
    grad[W0:B0] = (cache["feat"].astype(np.float64).T @ dZ).reshape(-1)
    grad[B0:WL] = dZ.sum(axis=0)
    return grad


def fd_check(theta, g: Graph, R, E, idx, eps=1e-6):
    """Central finite differences of the f64 loss at parameter indices idx."""
    th = np.asarray(theta, dtype=np.float64).copy()
    out = []
    for i in idx:
        t0 = th[i]
        th[i] = t0 + eps
        lp = loss_and_grad(th, g, R, E, "f64")[0]
        th[i] = t0 - eps
        lm = loss_and_grad(th, g, R, E, "f64")[0]
        th[i] = t0
        out.append((lp - lm) / (2 * eps))
    return np.array(out)

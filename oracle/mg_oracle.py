"""CPU restatement of the aggregation multilevel preconditioner -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the device library
(mpi-cst-fem-solver_amd/csrc/mg.hip).  The reference solver has no multilevel
preconditioner; this module restates the device algorithm step by step:

  aggregation   greedy in node order over the node graph (rows of the block pattern, self
                included): a node whose neighbours are all unaggregated becomes a root and takes
                its neighbourhood; then every leftover joins the aggregate of its first
                aggregated neighbour in column order; nodes left after that start their own.
  prolongator   per aggregate, modified Gram-Schmidt on the near-null-space rows of its DOFs
                (member nodes ascending, components in order): P = Q, next B = R; a column
                whose remaining norm is <= 1e-10 of its original norm (or zero) is dead (zero).
                Level 0's B: (1, 0, -(y - yc)) for ux, (0, 1, x - xc) for uy, (xc, yc) the
                centre of the mesh's bounding box (global modes, so every level's R carries
                them); rows of constrained DOFs are zero.
  Galerkin      A_{l+1} = P^T A_l P.
  cycle         level 0: z = S(r); z += P K_1(P^T (r - A z)); z += S(r - A z), with S = the
                learned M (given as a matrix) or two Chebyshev steps on D^-1 A over [lam/30, lam];
                levels >= 1: two Chebyshev steps, coarse correction, two steps; the coarse
                correction of level l is the dense solve on the coarsest level, else, for
                levels l+1 <= kdepth, the K-cycle of level l+1: two flexible CG steps (x = 0,
                alpha = p'r/p'q with 0 when p'q <= 0, beta = -z'q/p'q) preconditioned by the
                cycle of level l+1, and below that one cycle of level l+1 (V-cycle).
  fcg           the flexible CG the preconditioner runs in (FCG(1), the device's recurrence).

Pinning (tests/test_oracle_mg.py): known answers for the aggregation and the Gram-Schmidt
step; the prolongator is orthonormal per aggregate and reproduces the near-null space
(P R = B); coarse matrices are symmetric and the global rigid-body modes of the unconstrained
operator stay in the range of every prolongator and in the null space of every coarse matrix;
the two-level cycle is symmetric positive; CG with the cycle needs a small fraction of
Jacobi-CG's iterations; and fcg with the Jacobi preconditioner reproduces the C oracle's
Jacobi-PCG (the same Krylov space: iterations and solution).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def aggregate(ptr, col):
    n = len(ptr) - 1
    agg = -np.ones(n, dtype=np.int64)
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        nb = col[ptr[i]:ptr[i + 1]]
        if np.all(agg[nb] < 0):
            agg[nb] = na
            agg[i] = na
            na += 1
    for i in range(n):
        if agg[i] >= 0:
            continue
        for j in col[ptr[i]:ptr[i + 1]]:
            if agg[j] >= 0:
                agg[i] = agg[j]
                break
    for i in range(n):
        if agg[i] < 0:
            agg[i] = na
            na += 1
    return agg, na


def members(agg, na):
    order = np.argsort(agg, kind="stable")
    ptr = np.zeros(na + 1, dtype=np.int64)
    np.add.at(ptr, agg + 1, 1)
    return np.cumsum(ptr), order


def mgs(Bl):
    """Modified Gram-Schmidt of the rows Bl (m x 3): (Q m x 3, R 3 x 3), dead columns zero."""
    m = Bl.shape[0]
    q = np.array(Bl, dtype=np.float64).T.copy()  # 3 x m
    R = np.zeros((3, 3))
    alive = [False, False, False]
    for c in range(3):
        n0 = np.sqrt(np.sum(q[c] * q[c]))
        for cc in range(c):
            if not alive[cc]:
                continue
            dot = np.sum(q[cc] * q[c])
            R[cc, c] = dot
            q[c] = q[c] - dot * q[cc]
        nr = np.sqrt(np.sum(q[c] * q[c]))
        if n0 > 0 and nr > 1e-10 * n0:
            alive[c] = True
            R[c, c] = nr
            q[c] = q[c] / nr
        else:
            R[c, :] = 0.0
            q[c] = 0.0
    for c in range(3):
        if not alive[c]:
            R[c, :] = 0.0
    return q.T, R


class Level:
    pass


def block_pattern(A, bs):
    """Node graph (ptr, col) of a scalar CSR matrix whose DOFs come in blocks of bs."""
    n = A.shape[0] // bs
    C = sp.csr_matrix((np.ones(A.nnz), A.indices // bs, A.indptr), shape=(A.shape[0], n))[0::bs]
    C = C.tocsr()
    C.sum_duplicates()
    C.sort_indices()
    return C.indptr.astype(np.int64), C.indices.astype(np.int64)


def setup(A0, xy, fixed, max_levels=10, coarse_max=500, ptr0=None, col0=None):
    """Hierarchy of scalar CSR matrices; level 0 = A0 (2 DOFs per node)."""
    levels = []
    A, bs, B = A0.tocsr(), 2, None
    for l in range(max_levels):
        L = Level()
        L.A, L.bs, L.n = A, bs, A.shape[0]
        if l == 0 and ptr0 is not None:
            L.ptr, L.col = np.asarray(ptr0, np.int64), np.asarray(col0, np.int64)
        else:
            L.ptr, L.col = block_pattern(A, bs)
        levels.append(L)
        last = l + 1 >= max_levels - 1
        L.agg, L.na = aggregate(L.ptr, L.col)
        mptr, mem = members(L.agg, L.na)
        L.mptr, L.mem = mptr, mem
        if l == 0:
            cx = 0.5 * (xy[:, 0].min() + xy[:, 0].max())
            cy = 0.5 * (xy[:, 1].min() + xy[:, 1].max())
            dx = xy[:, 0] - cx
            dy = xy[:, 1] - cy
            B = np.zeros((L.n, 3))
            B[0::2, 0] = 1.0
            B[0::2, 2] = -dy
            B[1::2, 1] = 1.0
            B[1::2, 2] = dx
            B[np.asarray(fixed, bool)] = 0.0
        P = np.zeros((L.n, 3))
        Bn = np.zeros((3 * L.na, 3))
        for a in range(L.na):
            nodes = mem[mptr[a]:mptr[a + 1]]
            rows = (bs * nodes[:, None] + np.arange(bs)).reshape(-1)
            q, r = mgs(B[rows])
            P[rows] = q
            Bn[3 * a:3 * a + 3] = r
        dof_agg = np.repeat(L.agg, bs)
        L.Pvals = P
        L.P = sp.csr_matrix((P.reshape(-1), (np.repeat(np.arange(L.n), 3),
                                             (3 * dof_agg[:, None] + np.arange(3)).reshape(-1))),
                            shape=(L.n, 3 * L.na))
        L.B = B
        A = (L.P.T @ L.A @ L.P).tocsr()
        B, bs = Bn, 3
        if last or A.shape[0] <= coarse_max or L.na == len(L.ptr) - 1:
            C = Level()
            C.A, C.bs, C.n = A, 3, A.shape[0]
            levels.append(C)
            break
    Ac = levels[-1].A.toarray()
    dead = ~(np.diag(Ac) > 0)
    Ac[dead, :] = 0.0
    Ac[:, dead] = 0.0
    Ac[dead, dead] = 1.0
    inv = np.linalg.inv(Ac)
    inv[dead, :] = 0.0
    inv[:, dead] = 0.0
    levels[-1].inv = inv
    for L in levels:
        d = L.A.diagonal()
        L.dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    return levels


def cheb_coeffs(lam):
    hi = lam
    lo = hi / 30.0
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    rho0 = 1.0 / sigma
    rho1 = 1.0 / (2.0 * sigma - rho0)
    return 1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta


def rowsum_matrix(L):
    """The matrix the cycle's row sums read on a level: L.As when a test sets it (values
    rounded as the device stores them), else L.A.  Galerkin products, diagonals and the
    coarsest inverse always come from L.A."""
    return getattr(L, "As", None) if getattr(L, "As", None) is not None else L.A


def cheb2(L, x, r, lam):
    c1r, c2d, c2r = cheb_coeffs(lam)
    A = rowsum_matrix(L)
    d = c1r * (L.dinv * (r - (A @ x if x is not None else 0.0)))
    x = (x if x is not None else 0.0) + d
    d = c2d * d + c2r * (L.dinv * (r - A @ x))
    return x + d


def coarse_correction(levels, l, rc, lams, kdepth=99):
    if l + 1 == len(levels) - 1:
        return levels[-1].inv @ rc
    if l + 1 <= kdepth:
        return kcycle(levels, l + 1, rc, lams, kdepth)
    return cycle(levels, l + 1, rc, lams, kdepth=kdepth)


def cycle(levels, l, r, lams, M=None, kdepth=99):
    L = levels[l]
    A = rowsum_matrix(L)
    if l == 0 and M is not None:
        x = M @ r
    else:
        x = cheb2(L, None, r, lams[l])
    rc = L.P.T @ (r - A @ x)
    x = x + L.P @ coarse_correction(levels, l, rc, lams, kdepth)
    if l == 0 and M is not None:
        return x + M @ (r - A @ x)
    return cheb2(L, x, r, lams[l])


def kcycle(levels, l, rhs, lams, kdepth=99):
    A = rowsum_matrix(levels[l])
    res = rhs.copy()
    z = cycle(levels, l, res, lams, kdepth=kdepth)
    p = z.copy()
    q = A @ p
    pq, pr = p @ q, p @ res
    alpha = pr / pq if pq > 0 else 0.0
    x = alpha * p
    res = res - alpha * q
    z = cycle(levels, l, res, lams, kdepth=kdepth)
    zq = z @ q
    beta = -zq / pq if pq > 0 else 0.0
    q = A @ z + beta * q
    p = z + beta * p
    pq, pr = p @ q, p @ res
    alpha = pr / pq if pq > 0 else 0.0
    return x + alpha * p


def fcg(A, b, precond, rtol=1e-10, max_iter=1000):
    """Flexible CG with one kept direction (Notay's FCG(1)), the device's recurrence
    (csrc/pcg.hip fcg_run): x = 0, r = b, z = B r, p = z; per iteration q = A p,
    alpha = p'r / p'q, x += alpha p, r -= alpha q, stop when ||r|| / ||b|| < rtol,
    z = B r, beta = -z'q / p'q, p = z + beta p.  Returns (x, iterations, residual history)."""
    x = np.zeros_like(b)
    r = b.copy()
    bnorm = np.linalg.norm(b)
    hist = [1.0]
    if bnorm == 0.0:
        return x, 0, hist
    p = precond(r)
    for k in range(1, max_iter + 1):
        q = A @ p
        pq = p @ q
        alpha = (p @ r) / pq
        x = x + alpha * p
        r = r - alpha * q
        hist.append(np.linalg.norm(r) / bnorm)
        if hist[-1] < rtol:
            return x, k, hist
        z = precond(r)
        beta = -(z @ q) / pq
        p = z + beta * p
    return x, max_iter, hist

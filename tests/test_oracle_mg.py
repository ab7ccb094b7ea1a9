"""CPU tests pinning the multilevel-preconditioner oracle (oracle/mg_oracle.py).

No reference counterpart exists; the oracle is pinned by known answers (aggregation of a path
graph, Gram-Schmidt of a single node's rigid-body modes) and by mathematics: orthonormal
prolongator columns that reproduce the near-null space (P R = B), coarse matrices that are
symmetric and keep the rigid-body modes in the null space of the unconstrained operator, a
symmetric two-level cycle, and FCG iteration counts far below Jacobi-CG's.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import cstfem_hip
from cstfem_hip import pipeline
from oracle import mg_oracle as mg


def plate_system(oracle, n=12, constrained=True):
    prob = pipeline.plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=2, E=1.0, nu=0.3,
                                  plane=cstfem_hip.PLANE_STRAIN, tx=1.0, E_sigma=1.0,
                                  E_seed=22, body=(0.0, -1.0))
    m = prob.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=1.0, nu=0.3, strain=1, t=1.0, E_elem=prob.E_elem)
    f = np.zeros(m.num_dofs)
    for g, t in prob.tractions:
        oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=1.0, **t)
    oracle.add_body_force(m.xy, m.tri, f, 1.0, 0.0, -1.0)
    fixed = np.zeros(m.num_dofs, bool)
    if constrained:
        d, val = prob.dirichlet_dofs()
        K, b = oracle.apply_dirichlet(rp, c, v, f, d, val)
        fixed[d] = True
    else:
        K, b = v, f
    A = sp.csr_matrix((K, c, rp), shape=(m.num_dofs,) * 2)
    return A, b, m.xy, fixed


def test_aggregation_path_graph_kat():
    # path 0-1-2-3-4 (self loops included): 0 takes {0, 1}; 2 is blocked by 1; 3 takes {2, 3, 4}
    ptr = np.array([0, 2, 5, 8, 11, 13])
    col = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4])
    agg, na = mg.aggregate(ptr, col)
    assert na == 2 and list(agg) == [0, 0, 1, 1, 1]
    # a leftover joins its first aggregated neighbour: star 0-{1,2}, 3-{1}
    ptr = np.array([0, 3, 6, 8, 10])
    col = np.array([0, 1, 2, 0, 1, 3, 0, 2, 1, 3])
    agg, na = mg.aggregate(ptr, col)
    assert na == 1 and list(agg) == [0, 0, 0, 0]


def test_mgs_single_node_drops_rotation():
    B = np.array([[1.0, 0.0, -0.0], [0.0, 1.0, 0.0]])   # one node at the centroid
    Q, R = mg.mgs(B)
    np.testing.assert_array_equal(Q, [[1, 0, 0], [0, 1, 0]])
    np.testing.assert_array_equal(R, [[1, 0, 0], [0, 1, 0], [0, 0, 0]])


def test_prolongator_orthonormal_and_reproduces_modes(oracle):
    A, b, xy, fixed = plate_system(oracle)
    levels = mg.setup(A, xy, fixed, max_levels=3, coarse_max=20)
    for l, L in enumerate(levels[:-1]):
        PtP = (L.P.T @ L.P).toarray()
        alive = np.diag(PtP) > 0.5
        np.testing.assert_allclose(PtP, np.diag(alive.astype(float)), atol=1e-13)
        # P R = B on every aggregate (R = the next level's near-null space)
        Pd = L.Pvals
        for a in range(L.na):
            nodes = L.mem[L.mptr[a]:L.mptr[a + 1]]
            rows = (L.bs * nodes[:, None] + np.arange(L.bs)).reshape(-1)
            q, r = mg.mgs(L.B[rows])
            np.testing.assert_allclose(Pd[rows] @ r, L.B[rows], atol=1e-12 * max(1.0, np.abs(L.B[rows]).max()))


def test_coarse_matrices_symmetric_and_keep_rigid_modes(oracle):
    A, b, xy, fixed = plate_system(oracle, constrained=False)
    levels = mg.setup(A, xy, fixed, max_levels=4, coarse_max=20)
    assert len(levels) >= 3
    L0 = levels[0]
    # the global rigid-body modes are in the null space of the unconstrained stiffness ...
    G = np.zeros((A.shape[0], 3))
    G[0::2, 0] = 1.0
    G[1::2, 1] = 1.0
    G[0::2, 2] = -xy[:, 1]
    G[1::2, 2] = xy[:, 0]
    assert np.abs(A @ G).max() < 1e-10 * abs(A).max() * np.abs(G).max()
    # ... lie in the range of P (each aggregate carries translations and a rotation) ...
    Gc = L0.P.T @ G
    np.testing.assert_allclose(L0.P @ Gc, G, atol=1e-12 * np.abs(G).max())
    # ... so their coarse representation is in the coarse null space; A1 is symmetric
    A1 = levels[1].A
    assert abs(A1 - A1.T).max() < 1e-13 * abs(A1).max()
    assert np.abs(A1 @ Gc).max() < 1e-10 * abs(A1).max() * np.abs(Gc).max()
    # and the next level's prolongator (built from the R factors) still reproduces them, so
    # the global modes survive to level 2 (a rotation about per-aggregate centres would not)
    L1 = levels[1]
    Gc2 = L1.P.T @ Gc
    np.testing.assert_allclose(L1.P @ Gc2, Gc, atol=1e-10 * np.abs(Gc).max())
    A2 = levels[2].A
    assert np.abs(A2 @ Gc2).max() < 1e-9 * abs(A2).max() * np.abs(Gc2).max()


def test_two_level_cycle_is_symmetric(oracle):
    A, b, xy, fixed = plate_system(oracle)
    levels = mg.setup(A, xy, fixed, max_levels=2, coarse_max=10 ** 6)
    assert len(levels) == 2
    lams = [3.9, 0.0]
    rng = np.random.default_rng(0)
    r1, r2 = rng.standard_normal((2, A.shape[0]))
    z1 = mg.cycle(levels, 0, r1, lams)
    z2 = mg.cycle(levels, 0, r2, lams)
    assert abs(r2 @ z1 - r1 @ z2) < 1e-10 * abs(r2 @ z1)
    # and positive
    assert r1 @ z1 > 0 and r2 @ z2 > 0


def test_multilevel_fcg_beats_jacobi(oracle):
    A, b, xy, fixed = plate_system(oracle, n=40)
    levels = mg.setup(A, xy, fixed, max_levels=4, coarse_max=60)
    assert len(levels) >= 3
    lams = [float(abs(sla.eigsh(sp.diags(L.dinv) @ L.A, k=1, which="LM",
                                return_eigenvectors=False)[0])) * 1.1 for L in levels]
    counts = {}
    for name, Mf in (("jacobi", lambda r: levels[0].dinv * r),
                     ("multilevel", lambda r: mg.cycle(levels, 0, r, lams))):
        it = [0]

        def cb(xk):
            it[0] += 1
        x, info = sla.cg(A, b, rtol=1e-10, maxiter=5000, M=sla.LinearOperator(A.shape, Mf),
                         callback=cb)
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9
        counts[name] = it[0]
    assert counts["multilevel"] * 8 < counts["jacobi"], counts


def test_fcg_with_jacobi_reproduces_pcg(oracle):
    """FCG(1) with a fixed SPD preconditioner spans the same Krylov space as PCG, so with
    B = D^-1 it reproduces the C oracle's Jacobi-PCG (solver.cpp:35-105): iterations within
    one, solution to 1e-8; a dropped z'q term (beta = 0, steepest descent) does not."""
    A, b, _, _ = plate_system(oracle, n=24)
    dinv = 1.0 / A.diagonal()
    x, its, hist = mg.fcg(A, b, lambda r: dinv * r, rtol=1e-10, max_iter=5000)
    rp, c, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    x_ref, st = oracle.pcg_jacobi(rp, c, v, b, 1e-10, 5000)
    assert abs(its - st["iterations"]) <= 1, (its, st["iterations"])
    assert np.max(np.abs(x - x_ref)) <= 1e-8 * np.max(np.abs(x_ref))
    assert len(hist) == its + 1 and hist[-1] < 1e-10

    def steepest(A, b, precond, rtol, max_iter):  # FCG with the z'q term planted as zero
        x, r = np.zeros_like(b), b.copy()
        for k in range(1, max_iter + 1):
            p = precond(r)
            q = A @ p
            x, r = x + (p @ r) / (p @ q) * p, r - (p @ r) / (p @ q) * q
            if np.linalg.norm(r) < rtol * np.linalg.norm(b):
                return k
        return max_iter
    assert steepest(A, b, lambda r: dinv * r, 1e-10, 5000) > 3 * its

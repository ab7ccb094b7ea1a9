"""The learned preconditioner's training protocol at a reduced size: plates generated and
solved in batches, pairs harvested from every plate's plain Jacobi-CG iterates, the network
trained on them (plus white-noise pairs) and deployed as the multilevel flexible CG's fine
smoother on a heterogeneous plate of the same family.

- the harvested pairs of a sample plate equal the oracle's: e_k = x* - x_k with x_k the
  oracle's Jacobi-PCG iterate after k iterations (src/solver.cpp:35-105) and x* its solution,
  within FP32 rounding and the solve tolerance; r_k = K e_k;
- projecting the pairs onto the coarse complement keeps r = K e and lowers the energy of
  the (smooth) early errors; white-noise pairs drawn on the device are N(0, 1)
  and leave the earlier pairs in place;
- the trained network, deployed, reaches an FP64 true relative residual below 1e-10
  (recomputed on the host).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cstfem_hip
from cstfem_hip import learned, pipeline

pytestmark = pytest.mark.gpu


def oracle_system(oracle, p):
    m = p.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=1.0, nu=0.3, strain=1, t=1.0, E_elem=p.E_elem)
    f = np.zeros(m.num_dofs)
    for g, kw in p.tractions:
        oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=1.0, **kw)
    d, val = p.dirichlet_dofs()
    K, b = oracle.apply_dirichlet(rp, c, v, f, d, val)
    return rp, c, K, b


def test_batch_harvest_matches_oracle_iterates(dev, oracle):
    bp = pipeline.harvest_batch(count=6, seed=4, cells=20)
    pipeline.load_batch(dev, bp)
    dev.mpnn_init(0)
    K = dev.batch_harvest(1, 12)
    assert K == 12
    n = dev.n
    R = dev.mpnn_download(4, K * n).reshape(K, n)
    E = dev.mpnn_download(5, K * n).reshape(K, n)
    for k_plate in (0, 5):
        p = bp.plates[k_plate]
        o, m = 2 * int(bp.node_off[k_plate]), p.mesh.num_dofs
        rp, c, Kv, b = oracle_system(oracle, p)
        xs, st = oracle.pcg_jacobi(rp, c, Kv, b, 1e-12, 100000)
        A = sp.csr_matrix((Kv, c, rp), shape=(m, m))
        for k in (1, 5, 12):
            xk, _ = oracle.pcg_jacobi(rp, c, Kv, b, 1e-300, k)
            e = xs - xk
            ed = E[k - 1, o:o + m].astype(np.float64)
            assert np.max(np.abs(ed - e)) <= 1e-6 * np.max(np.abs(e)), (k_plate, k)
            # r_k = K e_k formed in FP64 on the device, then rounded to FP32
            rd = R[k - 1, o:o + m].astype(np.float64)
            re = A @ e
            assert np.max(np.abs(rd - re)) <= 1e-5 * np.max(np.abs(re)), (k_plate, k)


def test_projected_pairs_leave_coarse_correction_nothing(dev):
    bp = pipeline.harvest_batch(count=4, seed=9, cells=24)
    pipeline.load_batch(dev, bp)
    dev.mpnn_init(0)
    K = dev.batch_harvest(1, 4)
    E0 = dev.mpnn_download(5, K * dev.n).reshape(K, dev.n).astype(np.float64)
    dev.mg_setup(8, 200, 1, 3)
    dev.mpnn_project_pairs()
    E = dev.mpnn_download(5, K * dev.n).reshape(K, dev.n).astype(np.float64)
    R = dev.mpnn_download(4, K * dev.n).reshape(K, dev.n).astype(np.float64)
    A = dev.spmv
    for k in range(K):
        # r' = K e'; the coarse correction (a near K-orthogonal projection onto the coarse
        # space) takes energy away from the smooth early errors: |e'|_K < |e|_K
        assert np.linalg.norm(R[k] - A(E[k])) <= 1e-5 * np.linalg.norm(R[k])
        ratio = (E[k] @ A(E[k])) / (E0[k] @ A(E0[k]))
        print(f"pair {k}: energy after / before the coarse projection {ratio:.3f}")
        assert ratio < 0.95, k
    dev.mpnn_init(0)
    dev.mpnn_add_white_pairs(3, 11)
    E2 = dev.mpnn_download(5, (K + 3) * dev.n).reshape(K + 3, dev.n)
    assert np.array_equal(E2[:K].astype(np.float64), E)  # earlier pairs kept
    w = E2[K:].astype(np.float64)
    assert abs(w.mean()) < 0.02 and abs(w.std() - 1.0) < 0.02  # N(0, 1)


def test_protocol_trains_and_deploys(dev):
    theta, losses, info = learned.train_harvested(dev, plates=24, chunk=12, cells=24, k_first=1,
                                                  k_last=20, steps_per_chunk=15, white=8)
    assert np.all(np.isfinite(losses)) and info["pairs"] == 24 * 20
    prob = pipeline.plate_problem(96, 96, Lx=2.0, Ly=1.0, jitter=0.2, seed=3,
                                  plane=cstfem_hip.PLANE_STRAIN, E_sigma=1.0, E_seed=5)
    pipeline.load_problem(dev, prob)
    pipeline.assemble_and_constrain(dev, prob, multilevel=True)
    dev.setup()
    dev.mpnn_set_params(theta)
    dev.mpnn_build()
    dev.mg_setup(10, 300, 0, 3)
    x, st = dev.solve(rtol=1e-10, max_iter=2000, method=cstfem_hip.METHOD_FCG_LEARNED)
    print(f"protocol network: lambda_max(MK) {dev.mg_level_info(0)[3]:.3f}, {st.message}, "
          f"fallback {st.fallback}")
    b = dev.download_rhs()
    assert st.converged and np.linalg.norm(b - dev.spmv(x)) < 1.05e-10 * np.linalg.norm(b)

"""The batched workload at its benchmark size: 512 plates (seed 3, 16..160 cells per side,
8.3M DOFs) in one persistent launch per method.

- every plate's FP64 true residual ||b - K x|| / ||b||, recomputed on the host from the
  downloaded K, b and x, is below 1e-10 (the device's own check) up to the summation-order
  noise of the host's recomputation (5%), for FP64 and for FP32 with FP64 refinement;
- for a fixed sample (every 16th plate and the largest), the plate's block of the batched
  pattern, K and b equal the oracle's single-plate arrays bit for bit, and the solution
  matches the oracle's single-plate Jacobi-CG (src/solver.cpp:35-131): iterations +-2,
  solution within 1e-8 relative;
- the team launch (the largest plates split over up to 16 co-resident workgroups, the
  default) and one workgroup per plate give the same bits, in FP64 and in FP32.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cstfem_hip
from cstfem_hip import pipeline

from test_gpu_batched import oracle_plate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch512():
    return pipeline.batch_problem(count=512, seed=3, nmin=16, nmax=160)


def host_matrix(dev):
    rp, c = dev.download_pattern()
    K = dev.download_values()
    return rp, c, K, sp.csr_matrix((K, c, rp), shape=(dev.n, dev.n))


def per_plate_true_residual(bp, A, b, x):
    r = b - A @ x
    out = np.empty(bp.count)
    for k, p in enumerate(bp.plates):
        o, n = 2 * int(bp.node_off[k]), p.mesh.num_dofs
        out[k] = np.linalg.norm(r[o:o + n]) / np.linalg.norm(b[o:o + n])
    return out


def sample(bp):
    sizes = [p.mesh.num_dofs for p in bp.plates]
    return sorted(set(range(0, bp.count, 16)) | {int(np.argmax(sizes))})


def test_batch512_fp64_vs_oracle(dev, oracle, batch512):
    bp = batch512
    pipeline.load_batch(dev, bp)
    rp_all, c_all, K_all, A = host_matrix(dev)
    b_all = dev.download_rhs()
    x, res, secs = dev.batch_solve(method=cstfem_hip.METHOD_CG_FP64_REFINED)
    print(f"{bp.count} plates, {dev.n} DOFs, FP64 batch in {secs:.3f} s")
    tr = per_plate_true_residual(bp, A, b_all, x)
    assert all(r.converged for r in res)
    assert np.all(tr < 1.05e-10), (int(np.argmax(tr)), float(tr.max()))
    for k in sample(bp):
        p = bp.plates[k]
        rp, c, K, b, xo, st = oracle_plate(oracle, p)
        o, n = int(bp.node_off[k]), p.mesh.num_dofs
        lo, hi = rp_all[2 * o], rp_all[2 * o + n]
        assert np.array_equal(rp_all[2 * o:2 * o + n + 1] - lo, rp), k
        assert np.array_equal(c_all[lo:hi] - 2 * o, c), k
        assert np.array_equal(K_all[lo:hi], K), k
        assert np.array_equal(b_all[2 * o:2 * o + n], b), k
        assert abs(res[k].iterations - st["iterations"]) <= 2, (k, res[k].iterations,
                                                               st["iterations"])
        xs = x[2 * o:2 * o + n]
        assert np.max(np.abs(xs - xo)) <= 1e-8 * np.max(np.abs(xo)), k


def test_batch512_fp32_refined_true_residuals(dev, batch512):
    bp = batch512
    pipeline.load_batch(dev, bp)
    _, _, _, A = host_matrix(dev)
    b = dev.download_rhs()
    x32, r32, secs = dev.batch_solve(method=cstfem_hip.METHOD_IR_FP32)
    print(f"FP32 + FP64 refinement batch in {secs:.3f} s")
    assert all(r.converged for r in r32)
    tr = per_plate_true_residual(bp, A, b, x32)
    assert np.all(tr < 1.05e-10), (int(np.argmax(tr)), float(tr.max()))


@pytest.mark.parametrize("method", [cstfem_hip.METHOD_CG_FP64_REFINED, cstfem_hip.METHOD_IR_FP32])
def test_batch512_teams_are_bit_identical(dev, batch512, method):
    bp = batch512
    pipeline.load_batch(dev, bp)
    x1, r1, t1 = dev.batch_solve(method=method, config=(16, 1, 1))
    x16, r16, t16 = dev.batch_solve(method=method, config=(16, 1, 16))
    print(f"method {method}: one workgroup per plate {t1:.3f} s, teams of up to 16 {t16:.3f} s")
    assert np.array_equal(x1, x16)
    assert [r.iterations for r in r1] == [r.iterations for r in r16]
    assert all(r.converged for r in r16)

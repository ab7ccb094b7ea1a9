"""Batched persistent Jacobi-CG over independent plates: every plate equals its own
single-plate solve (oracle), the assembled block of every plate is bitwise the oracle's, the
launch shape (workgroup size, workgroups per CU, teams of workgroups per plate) never changes
results, and the FP32 mode reaches the FP64-verified tolerance."""
import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import pipeline

pytestmark = pytest.mark.gpu


def oracle_plate(oracle, p, rtol=1e-10):
    m = p.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=p.E, nu=p.nu, strain=0, t=p.thickness)
    f = np.zeros(m.num_dofs)
    for g, kw in p.tractions:
        oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=p.thickness, **kw)
    d, val = p.dirichlet_dofs()
    K, b = oracle.apply_dirichlet(rp, c, v, f, d, val)
    x, st = oracle.pcg_jacobi(rp, c, K, b, rtol, 100000)
    return rp, c, K, b, x, st


@pytest.fixture(scope="module")
def small_batch():
    return pipeline.batch_problem(count=14, seed=5, nmin=3, nmax=28)


def test_batch_matches_single_plates(dev, oracle, small_batch):
    bp = small_batch
    pipeline.load_batch(dev, bp)
    K_all = dev.download_values()
    b_all = dev.download_rhs()
    rp_all, c_all = dev.download_pattern()
    x, res, secs = dev.batch_solve(method=cstfem_hip.METHOD_CG_FP64_REFINED)
    print(f"{bp.count} plates, {dev.n} DOFs in {secs * 1e3:.2f} ms")
    for k, p in enumerate(bp.plates):
        rp, c, K, b, xo, st = oracle_plate(oracle, p)
        o = int(bp.node_off[k])
        n = p.mesh.num_dofs
        # the plate's rows of the batched CSR are the plate's CSR shifted by 2*o
        lo, hi = rp_all[2 * o], rp_all[2 * o + n]
        assert np.array_equal(rp_all[2 * o:2 * o + n + 1] - lo, rp)
        assert np.array_equal(c_all[lo:hi] - 2 * o, c)
        assert np.array_equal(K_all[lo:hi], K), k
        assert np.array_equal(b_all[2 * o:2 * o + n], b), k
        r = res[k]
        assert r.converged and r.true_rel_residual < 1e-10, (k, r.true_rel_residual)
        assert abs(r.iterations - st["iterations"]) <= 2, (k, r.iterations, st["iterations"])
        xs = x[2 * o:2 * o + n]
        assert np.max(np.abs(xs - xo)) <= 1e-8 * np.max(np.abs(xo)), k


def test_batch_launch_shape_invariance_and_autotune(dev, small_batch, tmp_path):
    """Launch shapes, one workgroup per plate (max_team 1) or teams of up to 2..16 (the
    largest plates split over co-resident workgroups), give the same bits: every reduction
    sums the same per-slice partials in one fixed order."""
    bp = small_batch
    pipeline.load_batch(dev, bp)
    x0, r0, _ = dev.batch_solve(config=(4, 1, 1))
    for cfg in ((16, 2), (4, 1, 16), (4, 2, 2), (8, 1, 4), (16, 1, 16)):
        x1, r1, _ = dev.batch_solve(config=cfg)
        assert np.array_equal(x0, x1), cfg
        assert [r.iterations for r in r0] == [r.iterations for r in r1], cfg
        assert [r.refinements for r in r0] == [r.refinements for r in r1], cfg
    cache = str(tmp_path / "autotune.txt")
    best, t = dev.batch_autotune(cache_path=cache)
    assert best[0] in (4, 8, 16) and best[1] in (1, 2, 4) and best[2] in (1, 16) and t > 0
    best2, _ = dev.batch_autotune(cache_path=cache)  # served from the cache
    assert best2 == best
    x2, r2, _ = dev.batch_solve(config=best)
    assert np.array_equal(x0, x2)


def test_team_barrier_gives_up_cleanly(dev, small_batch, monkeypatch):
    """A team barrier whose wait exceeds its bound flags the launch and every member of that
    team leaves its plate.  Forced from the host with a bound of one poll (nothing is wrong,
    the members just do not arrive within one poll of each other), the solve reports the
    failure instead of waiting; the next solve, with the default bound, gives the earlier
    bits again (the counters and the flag start at zero in every launch)."""
    bp = small_batch
    pipeline.load_batch(dev, bp)
    x0, r0, _ = dev.batch_solve(config=(4, 1, 16))
    monkeypatch.setenv("CSTFEM_BATCH_TEAM_SPIN", "1")
    with pytest.raises(cstfem_hip.CstfemError, match="team barrier timed out"):
        dev.batch_solve(config=(4, 1, 16))
    monkeypatch.delenv("CSTFEM_BATCH_TEAM_SPIN")
    x1, r1, _ = dev.batch_solve(config=(4, 1, 16))
    assert np.array_equal(x0, x1)
    assert [r.iterations for r in r0] == [r.iterations for r in r1]


def test_batch_fp32_refined(dev, small_batch):
    bp = small_batch
    pipeline.load_batch(dev, bp)
    x64, r64, t64 = dev.batch_solve(method=cstfem_hip.METHOD_CG_FP64_REFINED)
    x32, r32, t32 = dev.batch_solve(method=cstfem_hip.METHOD_IR_FP32)
    assert all(r.converged and r.true_rel_residual < 1e-10 for r in r32)
    # true residuals below 1e-10 bound the difference by the conditioning; loose bar
    assert np.max(np.abs(x32 - x64)) <= 1e-6 * np.max(np.abs(x64))
    b = dev.download_rhs()
    rr = b - dev.spmv(x32)
    for k, p in enumerate(bp.plates):
        o, n = int(bp.node_off[k]), p.mesh.num_dofs
        assert np.linalg.norm(rr[2 * o:2 * o + n]) < 1.01e-10 * np.linalg.norm(b[2 * o:2 * o + n])


def test_batch_rejects_coupled_plates(dev):
    prob = pipeline.plate_problem(6, 6)
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    with pytest.raises(cstfem_hip.CstfemError, match="shares an element"):
        dev.batch_define([0, 20, prob.mesh.num_nodes])

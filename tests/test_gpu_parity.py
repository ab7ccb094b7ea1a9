"""GPU parity of the HIP path against the oracle and the reference's golden fixtures.

Bars: pattern, assembled values, loads, Dirichlet-eliminated K and b, and the SpMV are
bit-exact (integer/index work and the reference's floating-point operation order); the
solution agrees within the solver tolerance and the iteration count within +-2 (only the
dot-product summation order differs, docs/BENCHMARKS.md:259 reports the same spread between
two CPU implementations).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import meshes, pipeline

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden():
    with open(os.path.join(GOLD, "golden.json")) as fh:
        return json.load(fh)


def kirsch_mesh(lvl):
    z = np.load(os.path.join(GOLD, f"kirsch_L{lvl}_mesh.npz"))
    groups = {k[6:]: z[k] for k in z.files if k.startswith("group_")}
    return meshes.Mesh(xy=z["xy"], tri=z["tri"], groups=groups)


def run_device(dev, prob, rtol=1e-10, max_iter=20000, **kw):
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    rp, c = dev.download_pattern()
    dev.assemble_stiffness()
    v = dev.download_values()
    dev.set_rhs(np.zeros(dev.n))
    for g, t in prob.tractions:
        dev.add_traction(prob.mesh.groups[g], **t)
    if prob.body is not None:
        dev.add_body_force(*prob.body)
    f = dev.download_rhs()
    d, val = prob.dirichlet_dofs()
    dev.apply_dirichlet(d, val)
    K = dev.download_values()
    b = dev.download_rhs()
    dev.setup()
    x, st = dev.solve(rtol=rtol, max_iter=max_iter, **kw)
    return dict(rp=rp, c=c, v=v, f=f, K=K, b=b, x=x, st=st)


def run_oracle(oracle, prob, rtol=1e-10, max_iter=20000):
    m = prob.mesh
    strain = 1 if prob.plane == cstfem_hip.PLANE_STRAIN else 0
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=prob.E, nu=prob.nu, strain=strain,
                        t=prob.thickness, E_elem=prob.E_elem)
    f = np.zeros(m.num_dofs)
    for g, t in prob.tractions:
        oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=prob.thickness, **t)
    if prob.body is not None:
        oracle.add_body_force(m.xy, m.tri, f, prob.thickness, *prob.body)
    d, val = prob.dirichlet_dofs()
    K, b = oracle.apply_dirichlet(rp, c, v, f, d, val)
    x, st = oracle.pcg_jacobi(rp, c, K, b, rtol, max_iter)
    return dict(rp=rp, c=c, v=v, f=f, K=K, b=b, x=x, st=st)


def rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("lvl", [0, 1])
def test_kirsch_golden(dev, lvl):
    """Reference Kirsch meshes (tests/data/kirsch_L{0,1}.msh) at rtol 1e-12: bitwise pattern,
    values and eliminated system; iterations 148 / 295 +-2 (verification_results.json)."""
    g = golden()[f"kirsch_L{lvl}"]
    m = kirsch_mesh(lvl)
    prob = pipeline.Problem(mesh=m, E=210e9, nu=0.3, thickness=0.01,
                            tractions=[("right", dict(kirsch=(100e6, 0.1))),
                                       ("top", dict(kirsch=(100e6, 0.1)))],
                            dirichlet=[("left_sym", 0, 0.0), ("bottom_sym", 1, 0.0)])
    r = run_device(dev, prob, rtol=1e-12)
    print(f"kirsch_L{lvl}: {r['st'].message}; true rel {r['st'].true_rel_residual:.3e}")
    assert sha(r["rp"]) == g["sha_row_ptr"] and sha(r["c"]) == g["sha_cols"]
    assert sha(r["v"]) == g["sha_vals"]
    assert sha(r["f"]) == g["sha_f"]
    assert sha(r["K"]) == g["sha_K"] and sha(r["b"]) == g["sha_b"]
    assert r["st"].converged
    assert abs(r["st"].iterations - g["iterations"]) <= 2
    x_ref = np.load(os.path.join(GOLD, f"kirsch_L{lvl}_x.npy"))
    assert rel_err(r["x"], x_ref) < 1e-9
    assert r["st"].true_rel_residual < 1e-11


def test_plate64_golden(dev):
    """Unit-square plate (64x64 cells): bitwise pattern/values/RHS against the reference."""
    g = golden()["plate64"]
    prob = pipeline.config_problem(0)
    r = run_device(dev, prob)
    print(f"plate64: {r['st'].message}")
    for key in ("row_ptr", "cols", "vals", "f", "K", "b"):
        arr = dict(row_ptr=r["rp"], cols=r["c"], vals=r["v"], f=r["f"], K=r["K"], b=r["b"])[key]
        assert sha(arr) == g[f"sha_{key}"], key
    assert abs(r["st"].iterations - g["iterations"]) <= 2
    x_ref = np.load(os.path.join(GOLD, "plate64_x.npy"))
    assert rel_err(r["x"], x_ref) < 1e-8
    assert r["st"].true_rel_residual < 1e-9


def test_kirsch_generator_golden(dev):
    g = golden()["kirsch40"]
    m = meshes.kirsch_quarter(40, 40, ratio=1.05, a=1.0, L=10.0)
    prob = pipeline.Problem(mesh=m, E=200e3, nu=0.3, tractions=[("right", dict(tx=1.0))],
                            dirichlet=[("left_sym", 0, 0.0), ("bottom_sym", 1, 0.0)])
    r = run_device(dev, prob)
    assert sha(r["rp"]) == g["sha_row_ptr"] and sha(r["c"]) == g["sha_cols"]
    assert sha(r["v"]) == g["sha_vals"] and sha(r["b"]) == g["sha_b"]
    assert abs(r["st"].iterations - g["iterations"]) <= 2


@pytest.mark.parametrize("strain,sigma,body", [(0, None, None), (1, 1.0, (0.0, -1.0)),
                                               (0, 0.5, (0.3, -1.0))])
def test_hetero_vs_oracle(dev, oracle, strain, sigma, body):
    """Per-element modulus and body force (project extensions): bitwise against the oracle."""
    prob = pipeline.plate_problem(48, 24, Lx=2.0, Ly=1.0, jitter=0.2, seed=5,
                                  plane=strain, E_sigma=sigma, E_seed=9, body=body)
    r = run_device(dev, prob)
    o = run_oracle(oracle, prob)
    for key in ("rp", "c", "v", "f", "K", "b"):
        assert np.array_equal(r[key], o[key]), key
    assert r["st"].converged and o["st"]["converged"]
    assert abs(r["st"].iterations - o["st"]["iterations"]) <= 2
    assert rel_err(r["x"], o["x"]) < 1e-8


def test_spmv_bitwise(dev, oracle):
    prob = pipeline.plate_problem(40, 30, jitter=0.2, seed=3, E_sigma=1.0)
    r = run_device(dev, prob, max_iter=1)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(dev.n)
    y = dev.spmv(x)
    y_ref = oracle.spmv(r["rp"], r["c"], r["K"], x)
    assert np.array_equal(y, y_ref)


def random_spd(n, rng):
    """test_pcg.cpp:57-73: M^T M + n I scaled by D with 10^(0..3) entries."""
    M = rng.uniform(-1, 1, (n, n))
    A = M.T @ M + n * np.eye(n)
    d = 10.0 ** (3.0 * (rng.uniform(-1, 1, n) + 1.0) / 2.0)
    return A * d[:, None] * d[None, :]


def dense_to_csr(A):
    n = A.shape[0]
    rp, c, v = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        c.extend(nz.tolist())
        v.extend(A[i, nz].tolist())
        rp.append(len(c))
    return np.array(rp, np.int32), np.array(c, np.int32), np.array(v)


@pytest.mark.parametrize("n", [5, 40, 120])
def test_random_spd_vs_cholesky(dev, n):
    """test_pcg.cpp:75-102 on the generic CSR upload (odd n included)."""
    rng = np.random.default_rng(2026 + n)
    A = random_spd(n, rng)
    b = rng.uniform(-1, 1, n)
    x_ref = np.linalg.solve(A, b)
    dev.upload_csr(*dense_to_csr(A))
    dev.set_rhs(b)
    dev.setup()
    x, st = dev.solve(rtol=1e-12, max_iter=1000, record_history=True)
    assert st.converged and st.rel_residual < 1e-12
    assert np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)) < 1e-9
    assert len(st.residual_history) == st.iterations + 1
    assert st.residual_history[0] == 1.0 and st.residual_history[-1] < 1e-12


def test_zero_rhs_not_converged_bad_diagonal(dev):
    rng = np.random.default_rng(1)
    A = random_spd(6, rng)
    dev.upload_csr(*dense_to_csr(A))
    dev.set_rhs(np.zeros(6))
    dev.setup()
    x, st = dev.solve()
    assert st.converged and st.iterations == 0 and np.all(x == 0.0)
    assert st.message == "zero right-hand side, x = 0"
    A = random_spd(30, np.random.default_rng(3))
    dev.upload_csr(*dense_to_csr(A))
    dev.set_rhs(np.ones(30))
    dev.setup()
    x, st = dev.solve(rtol=1e-14, max_iter=2)
    assert not st.converged and st.iterations == 2 and "NOT converged" in st.message
    # a NULL load vector is the zero vector, set on the device over the previous one
    dev.set_rhs(None)
    assert np.all(dev.download_rhs() == 0.0)
    dev.setup()
    x, st = dev.solve()
    assert st.converged and st.iterations == 0 and np.all(x == 0.0)
    dev.upload_csr(*dense_to_csr(np.array([[2.0, 1.0], [1.0, -1.0]])))
    with pytest.raises(cstfem_hip.CstfemError, match="non-positive diagonal"):
        dev.setup()


def test_symmetric_elimination_kat(dev):
    """test_bc.cpp:63-78: K = [4 1 0; 1 3 1; 0 1 2], b = 1, u_1 = 2 -> b = [-1, 6, -1],
    K = diag(4, 3, 2) with the zeros kept."""
    rp = np.array([0, 2, 5, 7], np.int32)
    c = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v = np.array([4, 1, 1, 3, 1, 1, 2], float)
    dev.upload_csr(rp, c, v)
    dev.set_rhs(np.ones(3))
    dev.apply_dirichlet([1], [2.0])
    assert np.array_equal(dev.download_rhs(), [-1.0, 6.0, -1.0])
    assert np.array_equal(dev.download_values(), [4, 0, 0, 3, 0, 0, 2])
    with pytest.raises(cstfem_hip.CstfemError, match="conflicting Dirichlet values for node 0"):
        dev.apply_dirichlet([1, 1], [2.0, 3.0])
    with pytest.raises(cstfem_hip.CstfemError, match="out of range"):
        dev.apply_dirichlet([7], [0.0])


def test_solve_before_setup(dev):
    dev.upload_csr(*dense_to_csr(np.eye(4) * 2))
    dev.set_rhs(np.ones(4))
    with pytest.raises(cstfem_hip.CstfemLogicError, match="before setup"):
        dev.solve()


def test_instrumented_bitwise(dev):
    """A timing-on run equals a timing-off run bit for bit (test_pcg.cpp:183-202)."""
    prob = pipeline.plate_problem(32, 32, seed=4)
    pipeline.load_problem(dev, prob)
    pipeline.assemble_and_constrain(dev, prob)
    dev.setup()
    x0, s0 = dev.solve(instrument=False)
    x1, s1 = dev.solve(instrument=True)
    assert s0.iterations == s1.iterations and np.array_equal(x0, x1)
    assert s1.breakdown["spmv_calls"] == s1.iterations + 1
    assert s1.breakdown["spmv_local"] > 0 and s1.breakdown["total"] > 0


def test_reproducible(dev):
    """Fixed-order reductions: two solves are bitwise identical."""
    prob = pipeline.plate_problem(96, 64, Lx=1.5, seed=6, E_sigma=1.0, body=(0.0, -1.0))
    pipeline.load_problem(dev, prob)
    pipeline.assemble_and_constrain(dev, prob)
    dev.setup()
    x0, s0 = dev.solve()
    x1, s1 = dev.solve()
    assert s0.iterations == s1.iterations and np.array_equal(x0, x1)


@pytest.mark.parametrize("method", [cstfem_hip.METHOD_CG_FP64_REFINED, cstfem_hip.METHOD_IR_FP32])
def test_refined_methods_reach_true_rtol(dev, oracle, method):
    """FP64-refined and FP32+FP64-refinement solves stop on the FP64 true residual; their
    solutions agree with the oracle's FP64 Jacobi-CG solution (rtol 1e-12) to 1e-8."""
    prob = pipeline.plate_problem(80, 40, Lx=2.0, Ly=1.0, jitter=0.2, seed=12,
                                  plane=cstfem_hip.PLANE_STRAIN, E_sigma=1.0, E_seed=3,
                                  body=(0.0, -1.0))
    o = run_oracle(oracle, prob, rtol=1e-13, max_iter=100000)
    pipeline.load_problem(dev, prob)
    pipeline.assemble_and_constrain(dev, prob)
    dev.setup()
    x, st = dev.solve(rtol=1e-10, max_iter=100000, method=method)
    b = dev.download_rhs()
    r = b - dev.spmv(x)
    true_rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"method {method}: {st.iterations} iterations, {st.outer_iterations} refinements, "
          f"true rel {st.true_rel_residual:.3e} (host check {true_rel:.3e}), "
          f"oracle {o['st']['iterations']} iterations; "
          f"max rel diff vs oracle {rel_err(x, o['x']):.2e}")
    assert st.converged and st.true_rel_residual < 1e-10
    assert true_rel < 1.5e-10
    assert rel_err(x, o["x"]) < 1e-7


def test_ir_fp32_zero_rhs_and_maxiter(dev):
    prob = pipeline.plate_problem(16, 16, seed=2)
    pipeline.load_problem(dev, prob)
    pipeline.assemble_and_constrain(dev, prob)
    dev.set_rhs(np.zeros(dev.n))
    dev.setup()
    x, st = dev.solve(method=cstfem_hip.METHOD_IR_FP32)
    assert st.converged and st.iterations == 0 and not np.any(x)
    pipeline.assemble_and_constrain(dev, prob)
    dev.setup()
    x, st = dev.solve(method=cstfem_hip.METHOD_IR_FP32, max_iter=4)
    assert not st.converged and st.iterations <= 4

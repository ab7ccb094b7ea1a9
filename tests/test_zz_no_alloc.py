"""The allocation-free Dirichlet elimination (cstfem_apply_dirichlet_noalloc) against the
allocating one (cstfem_apply_dirichlet), bit for bit, on two calls in a row, the second on a
smaller system; and its limits."""
import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import pipeline

pytestmark = pytest.mark.gpu


def constraints(prob, rng):
    """Edge constraints plus scattered interior DOFs, with non-zero prescribed values."""
    m = prob.mesh
    d = np.concatenate([2 * np.unique(m.groups["left"]), 2 * np.unique(m.groups["bottom"]) + 1])
    extra = rng.choice(m.num_dofs, max(8, m.num_dofs // 200), replace=False)
    d = np.concatenate([d, extra[~np.isin(extra, d)]]).astype(np.int32)
    return d, rng.uniform(-1.0, 1.0, d.size)


def eliminated(dev, prob, f, d, u, noalloc):
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    dev.assemble_stiffness()
    dev.set_rhs(f)
    (dev.apply_dirichlet_noalloc if noalloc else dev.apply_dirichlet)(d, u)
    return dev.download_values(), dev.download_rhs()


def test_noalloc_equals_allocating_twice(dev):
    rng = np.random.default_rng(2024)
    big = pipeline.plate_problem(240, 160, Lx=1.5, jitter=0.2, seed=4, E_sigma=1.0)
    small = pipeline.plate_problem(30, 20, jitter=0.2, seed=5)
    cases = []
    for prob in (big, small):
        d, u = constraints(prob, rng)
        f = rng.standard_normal(prob.mesh.num_dofs)
        cases.append((prob, f, d, u, eliminated(dev, prob, f, d, u, noalloc=False)))
    # two allocation-free calls in a row, the second smaller
    for prob, f, d, u, (K_ref, b_ref) in cases:
        K, b = eliminated(dev, prob, f, d, u, noalloc=True)
        assert np.array_equal(K, K_ref) and np.array_equal(b, b_ref), prob.mesh.num_dofs


def test_noalloc_nonsymmetric_pattern_and_errors(dev):
    rng = np.random.default_rng(31)
    n = 97
    A = np.where(rng.uniform(size=(n, n)) < 0.05, rng.standard_normal((n, n)), 0.0)
    A += np.diag(rng.uniform(1, 2, n) * n)
    rp = np.concatenate([[0], np.cumsum((A != 0).sum(axis=1))]).astype(np.int32)
    c = np.nonzero(A)[1].astype(np.int32)
    v = A[A != 0]
    b = rng.standard_normal(n)
    d = rng.choice(n, 11, replace=False).astype(np.int32)
    u = rng.uniform(-1, 1, 11)
    out = []
    for fn in ("apply_dirichlet", "apply_dirichlet_noalloc"):
        dev.upload_csr(rp, c, v)
        dev.set_rhs(b)
        getattr(dev, fn)(d, u)
        out.append((dev.download_values(), dev.download_rhs()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    # the reference's errors, unchanged
    with pytest.raises(cstfem_hip.CstfemError, match="conflicting Dirichlet values"):
        dev.apply_dirichlet_noalloc([3, 3], [1.0, 2.0])
    with pytest.raises(cstfem_hip.CstfemError, match="out of range"):
        dev.apply_dirichlet_noalloc([n], [0.0])


def test_noalloc_rejects_beyond_scratch(dev):
    """More than 2^18 distinct constrained DOFs: an error, and K and b unchanged."""
    prob = pipeline.plate_problem(400, 400, jitter=0.0, seed=1)
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    dev.assemble_stiffness()
    f = np.linspace(-1.0, 1.0, prob.mesh.num_dofs)
    dev.set_rhs(f)
    K0 = dev.download_values()
    d = np.arange(prob.mesh.num_dofs, dtype=np.int32)
    assert d.size > 2 ** 18
    with pytest.raises(cstfem_hip.CstfemError, match="scratch limit"):
        dev.apply_dirichlet_noalloc(d, np.zeros(d.size))
    assert np.array_equal(dev.download_values(), K0) and np.array_equal(dev.download_rhs(), f)
    # a call within the limits still works afterwards
    dev.apply_dirichlet_noalloc(d[:1000], np.zeros(1000))

"""CPU tests pinning the learned-preconditioner oracle (oracle/mpnn_oracle.py).

The reference has no learned preconditioner, so the oracle is pinned by mathematics:
central finite differences of its FP64 loss against its analytic backward (every parameter
block), the Jacobi identity of the initial decoder, symmetry of M for any parameters, a zero
loss on pairs built from M's own action, and BF16 rounding known answers.
"""
import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import pipeline
from oracle import mpnn_oracle as mo


def small_system(oracle, n=7, sigma=1.0):
    prob = pipeline.plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=2, E=1.0, nu=0.3,
                                  plane=cstfem_hip.PLANE_STRAIN, tx=1.0, E_sigma=sigma,
                                  E_seed=22, body=(0.0, -1.0))
    m = prob.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=prob.E, nu=prob.nu, strain=1, t=1.0,
                        E_elem=prob.E_elem)
    f = np.zeros(m.num_dofs)
    d, val = prob.dirichlet_dofs()
    K, _ = oracle.apply_dirichlet(rp, c, v, f, d, val)
    fixed = np.zeros(m.num_dofs)
    fixed[d] = 1.0
    g = mo.Graph(rp, c, K, m.tri, m.num_nodes, prob.E_elem, fixed)
    return prob, g


def random_params(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    th = mo.init_params(seed).astype(np.float64)
    th[mo.B0:mo.WL] = 0.1 * rng.standard_normal(mo.H)
    th[mo.BL:mo.WD] = 0.1 * rng.standard_normal(mo.ROUNDS * mo.H)
    th[mo.WD:mo.BD] = 0.02 * rng.standard_normal(mo.H * mo.DEC)
    th[mo.WO:mo.BO] = 0.02 * rng.standard_normal(mo.H * mo.DEC)
    th[mo.BD + 3] = -0.1
    th[mo.BO:] = [-0.05, 0.01, -0.05, -0.1]
    return th


def pairs(g, rng, K=3):
    E = rng.standard_normal((K, g.n))
    R = np.stack([g.A @ e for e in E])
    return R, E


def test_bf16_rounding_kats():
    x = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 1.0 + 2.0**-9, -2.5, 0.0,
                  np.inf, 3.0e38], dtype=np.float32)
    y = mo.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6); 1 + 2^-9 rounds down
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0 + 2.0**-6)
    assert y[3] == 1.0 and y[4] == -2.5 and y[5] == 0.0 and np.isinf(y[6])
    assert abs(y[7] - 3.0e38) / 3.0e38 < 2.0**-8


def test_initial_decoder_is_jacobi(oracle):
    _, g = small_system(oracle)
    th = mo.init_params(3)
    for prec in ("f64", "bf16"):
        ghat, _ = mo.forward(th, g, prec)
        M = mo.m_matrix(ghat, g).toarray()
        # sqrt(d)^2 may differ from d by an ulp; off-diagonal entries are exactly zero
        np.testing.assert_array_equal(M - np.diag(np.diag(M)), 0.0)
        np.testing.assert_allclose(np.diag(M), g.invd, rtol=4e-16, atol=0)


def test_polynomial_decoder(oracle):
    """Biases (alpha, 0, alpha, -beta) on both decoders, zero weights: M is the polynomial
    preconditioner D^-1 (alpha I - beta K D^-1), built here from the sparse matrix itself."""
    _, g = small_system(oracle)
    th = mo.init_params(3)
    th[mo.WD:mo.NPARAMS] = 0.0
    alpha, beta = 1.25, 0.375  # exact in FP32 (the parameter vector is FP32)
    th[mo.BD:mo.WO] = [alpha, 0.0, alpha, -beta]
    th[mo.BO:] = [0.0, 0.0, 0.0, -beta]
    ghat, _ = mo.forward(th, g, "f64")
    M = mo.m_matrix(ghat, g).toarray()
    Dinv = np.diag(g.invd)
    want = Dinv @ (alpha * np.eye(g.n) - beta * g.A.toarray() @ Dinv)
    np.testing.assert_allclose(M, want, rtol=1e-14, atol=1e-14 * np.abs(want).max())


def test_m_symmetric_any_params(oracle):
    _, g = small_system(oracle)
    ghat, _ = mo.forward(random_params(5), g, "f64")
    M = mo.m_matrix(ghat, g).toarray()
    assert np.abs(M).max() > 0
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-15 * np.abs(M).max())


def test_zero_loss_on_own_pairs(oracle):
    _, g = small_system(oracle)
    th = random_params(6)
    ghat, _ = mo.forward(th, g, "f64")
    M = mo.m_matrix(ghat, g)
    rng = np.random.Generator(np.random.PCG64(1))
    R = rng.standard_normal((2, g.n))
    E = np.stack([M @ r for r in R])
    loss, grad, _ = mo.loss_and_grad(th, g, R, E, "f64")
    assert abs(loss) < 1e-24 and np.abs(grad).max() < 1e-10


@pytest.mark.parametrize("block", ["W0", "B0", "WL0", "WL2", "BL1", "WD", "BD", "WO", "BO"])
def test_backward_vs_finite_differences(oracle, block):
    _, g = small_system(oracle)
    th = random_params(7)
    rng = np.random.Generator(np.random.PCG64(2))
    R, E = pairs(g, rng)
    # pairs whose loss is well away from zero
    E = E + 0.3 * rng.standard_normal(E.shape)
    loss, grad, _ = mo.loss_and_grad(th, g, R, E, "f64")
    lo, hi = {"W0": (mo.W0, mo.W0 + mo.LIVE * mo.H), "B0": (mo.B0, mo.WL),
              "WL0": (mo.WL, mo.WL + 2 * mo.H * mo.H),
              "WL2": (mo.WL + 4 * mo.H * mo.H, mo.BL), "BL1": (mo.BL + mo.H, mo.BL + 2 * mo.H),
              "WD": (mo.WD, mo.BD), "BD": (mo.BD, mo.WO), "WO": (mo.WO, mo.BO),
              "BO": (mo.BO, mo.NPARAMS)}[block]
    # the entries with the largest gradients in the block, plus random ones
    blk = np.arange(lo, hi)
    idx = np.unique(np.concatenate([blk[np.argsort(-np.abs(grad[lo:hi]))[:4]],
                                    rng.choice(blk, size=min(4, blk.size), replace=False)]))
    fd = mo.fd_check(th, g, R, E, idx, eps=1e-6)
    scale = max(np.abs(grad[lo:hi]).max(), 1e-12)
    assert np.abs(grad[lo:hi]).max() > 0
    np.testing.assert_allclose(grad[idx], fd, rtol=1e-5, atol=1e-6 * scale)


def test_features_scale_invariant(oracle):
    """Scaling E (hence K) by a constant leaves every feature unchanged."""
    _, g = small_system(oracle)
    f1 = g.features()
    g2 = mo.Graph(np.asarray(g.A.indptr), g.A.indices, g.A.data * 37.0, g.tri, g.nb,
                  g.E_elem * 37.0, g.fixed)
    np.testing.assert_allclose(g2.features(), f1, rtol=1e-12, atol=1e-12)


def test_package_layout_matches_oracle():
    """The package's parameter layout (cstfem_hip.learned) is the oracle's."""
    from cstfem_hip import learned
    for name in ("H", "F", "LIVE", "ROUNDS", "DEC", "W0", "B0", "WL", "BL", "WD", "BD", "WO",
                 "BO", "NPARAMS"):
        assert getattr(learned, name) == getattr(mo, name), name
    th = learned.polynomial_start(mo.init_params(0), beta=0.375)
    assert list(th[mo.BD:mo.WO]) == [1.0, 0.0, 1.0, -0.375]
    assert list(th[mo.BO:]) == [0.0, 0.0, 0.0, -0.375]
    assert not th[mo.WD:mo.BD].any() and not th[mo.WO:mo.BO].any()

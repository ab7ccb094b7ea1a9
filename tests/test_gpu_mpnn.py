"""Learned preconditioner on the GPU against the CPU oracle (oracle/mpnn_oracle.py).

- features, hidden state and decoder outputs of the BF16 MFMA network vs the oracle's BF16
  mode (same rounding points, FP32 accumulation in a different order);
- loss and gradient vs the oracle's BF16 mode (the oracle's backward is pinned by finite
  differences on the CPU);
- training on a fixed seed lowers the loss;
- flexible CG with M = Jacobi reproduces FP64 Jacobi-CG's iteration count;
- a NaN or an indefinite network triggers the Jacobi-CG fallback, which still converges;
- a trained network cuts the FP64 CG iterations on a plate it was not trained on.
No reference counterpart exists for this component.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cstfem_hip
from cstfem_hip import pipeline
from oracle import mpnn_oracle as mo

pytestmark = pytest.mark.gpu

FCG, FCG32 = cstfem_hip.METHOD_FCG_LEARNED, cstfem_hip.METHOD_FCG_LEARNED_FP32


def hetero(n, e_seed=22, seed=2):
    return pipeline.plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=seed, E=1.0, nu=0.3,
                                  plane=cstfem_hip.PLANE_STRAIN, tx=1.0, E_sigma=1.0,
                                  E_seed=e_seed, body=(0.0, -1.0))


def device_system(dev, prob):
    """Assembles, loads and constrains prob on the device; returns the oracle's Graph of the
    device matrix (the device values equal the oracle's bitwise, see test_gpu_parity)."""
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    rp, c = dev.download_pattern()
    dev.assemble_stiffness()
    dev.set_rhs(np.zeros(dev.n))
    for g, t in prob.tractions:
        dev.add_traction(prob.mesh.groups[g], **t)
    if prob.body is not None:
        dev.add_body_force(*prob.body)
    d, val = prob.dirichlet_dofs()
    dev.apply_dirichlet(d, val)
    K = dev.download_values()
    dev.setup()
    fixed = np.zeros(dev.n)
    fixed[d] = 1.0
    return mo.Graph(rp, c, K, prob.mesh.tri, prob.mesh.num_nodes, prob.E_elem, fixed)


def random_params(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    th = mo.init_params(seed)
    th[mo.B0:mo.WL] = 0.1 * rng.standard_normal(mo.H)
    th[mo.BL:mo.WD] = 0.1 * rng.standard_normal(mo.ROUNDS * mo.H)
    th[mo.WD:mo.BD] = 0.02 * rng.standard_normal(mo.H * mo.DEC)
    th[mo.WO:mo.BO] = 0.02 * rng.standard_normal(mo.H * mo.DEC)
    th[mo.BD + 3] = -0.1
    th[mo.BO:] = [-0.05, 0.01, -0.05, -0.1]
    return th.astype(np.float32)


def neumann_params(seed, beta=0.225):
    """Decoder at D^-1 (I - beta K D^-1), the degree-1 polynomial preconditioner."""
    th = mo.init_params(seed)
    th[mo.BD + 3] = -beta
    th[mo.BO + 3] = -beta
    return th


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def test_num_params(dev):
    assert dev.mpnn_num_params() == mo.NPARAMS


def test_forward_matches_oracle(dev):
    g = device_system(dev, hetero(24))
    th = random_params(7)
    dev.mpnn_init(0)
    dev.mpnn_set_params(th)
    dev.mpnn_build()
    ghat, cache = mo.forward(th, g, "bf16")
    feat = dev.mpnn_download(0, g.nb * mo.F).reshape(g.nb, mo.F)
    # the same FP64 feature values rounded through FP32 to BF16: at most one BF16 ulp apart
    assert np.all(np.abs(feat - cache["feat"]) <= 2.0**-7 * np.abs(cache["feat"]) + 1e-30)
    H = dev.mpnn_download(1, g.nb * mo.H).reshape(g.nb, mo.H)
    assert rel(H, cache["H"][-1]) < 1e-3
    gd = dev.mpnn_download(2, len(g.bcol) * mo.DEC).reshape(-1, mo.DEC)
    assert rel(gd, ghat) < 1e-3


def test_loss_and_gradient_match_oracle(dev):
    g = device_system(dev, hetero(24))
    th = random_params(7)
    dev.mpnn_init(0)
    dev.mpnn_set_params(th)
    rng = np.random.Generator(np.random.PCG64(1))
    E = rng.standard_normal((4, g.n))
    R = np.stack([g.A @ e for e in E]).astype(np.float32)
    E = (E + 0.3 * rng.standard_normal(E.shape)).astype(np.float32)
    dev.mpnn_set_pairs(R, E)
    loss = dev.mpnn_loss_grad()
    grad = dev.mpnn_download(3, mo.NPARAMS)
    lo, go, _ = mo.loss_and_grad(th, g, R.astype(np.float64), E.astype(np.float64), "bf16")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    for a, b in [(mo.W0, mo.B0), (mo.B0, mo.WL), (mo.WL, mo.BL), (mo.BL, mo.WD),
                 (mo.WD, mo.BD), (mo.BD, mo.WO), (mo.WO, mo.BO), (mo.BO, mo.NPARAMS)]:
        assert np.linalg.norm(go[a:b]) > 0
        assert rel(grad[a:b], go[a:b]) < 1e-3, (a, b)


def test_training_lowers_loss(dev):
    g = device_system(dev, hetero(32, e_seed=101, seed=5))
    dev.mpnn_init(1)
    dev.mpnn_set_params(neumann_params(1))
    rng = np.random.Generator(np.random.PCG64(3))
    E = rng.standard_normal((16, g.n))
    R = np.stack([g.A @ e for e in E])
    dev.mpnn_set_pairs(R.astype(np.float32), E.astype(np.float32))
    losses = [dev.mpnn_train_step(1e-3) for _ in range(120)]
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.9 * losses[0]
    # training is deterministic: the same steps from the same start give the same loss
    dev.mpnn_init(1)
    dev.mpnn_set_params(neumann_params(1))
    again = [dev.mpnn_train_step(1e-3) for _ in range(5)]
    assert again == losses[:5]


def test_fcg_with_jacobi_reproduces_cg(dev):
    device_system(dev, hetero(40))
    _, s_cg = dev.solve(rtol=1e-10, max_iter=20000, method=cstfem_hip.METHOD_CG_FP64)
    dev.mpnn_init(0)            # decoder = Jacobi
    dev.mpnn_build()
    x, s = dev.solve(rtol=1e-10, max_iter=20000, method=FCG)
    assert s.converged and not s.fallback
    assert abs(s.iterations - s_cg.iterations) <= 3
    assert s.true_rel_residual < 1e-10


@pytest.mark.parametrize("bad", ["nan", "indefinite"])
def test_fallback_on_bad_network(dev, bad):
    device_system(dev, hetero(24))
    _, s_ref = dev.solve(rtol=1e-10, max_iter=20000, method=cstfem_hip.METHOD_IR_FP32)
    th = mo.init_params(0)
    if bad == "nan":
        th[mo.BD] = np.nan          # M = NaN: breakdown at the first iteration
    else:
        th[mo.BD + 2] = -1.0        # M = diag(+D^-1, -D^-1): indefinite
    dev.mpnn_set_params(th)
    dev.mpnn_build()
    for method in (FCG, FCG32):
        x, s = dev.solve(rtol=1e-10, max_iter=20000, method=method)
        assert s.fallback == 1 and s.converged, (bad, method)  # Jacobi-CG
        assert s.true_rel_residual < 1e-10
        if bad == "nan":
            assert s.breakdown or s.iterations >= s_ref.iterations


def test_fcg_before_build_is_an_order_error(dev):
    device_system(dev, hetero(12))
    dev.mpnn_init(0)
    with pytest.raises(cstfem_hip.CstfemLogicError):
        dev.solve(rtol=1e-10, max_iter=100, method=FCG)


def test_trained_network_cuts_iterations(dev):
    """Train on one plate (white-noise error/residual pairs), solve another."""
    g = device_system(dev, hetero(48, e_seed=101, seed=5))
    dev.mpnn_init(1)
    dev.mpnn_set_params(neumann_params(1))
    rng = np.random.Generator(np.random.PCG64(3))
    E = rng.standard_normal((24, g.n))
    R = np.stack([g.A @ e for e in E])
    dev.mpnn_set_pairs(R.astype(np.float32), E.astype(np.float32))
    for _ in range(200):
        dev.mpnn_train_step(1e-3)
    theta = dev.mpnn_get_params()
    g2 = device_system(dev, hetero(80, e_seed=7, seed=9))
    _, s_cg = dev.solve(rtol=1e-10, max_iter=50000, method=cstfem_hip.METHOD_CG_FP64)
    dev.mpnn_set_params(theta)
    dev.mpnn_build()
    x, s = dev.solve(rtol=1e-10, max_iter=50000, method=FCG)
    assert s.converged and not s.fallback and s.true_rel_residual < 1e-10
    assert s.iterations < 0.65 * s_cg.iterations, (s.iterations, s_cg.iterations)
    # the solution of the system, whatever the preconditioner
    b = dev.download_rhs()
    assert np.linalg.norm(b - g2.A @ x) / np.linalg.norm(b) < 1e-10
    x32, s32 = dev.solve(rtol=1e-10, max_iter=50000, method=FCG32)
    assert s32.converged and s32.true_rel_residual < 1e-10


def test_harvest_pairs(dev):
    """Pairs from the plain solver's iterates: count, and a finite loss for the Jacobi M."""
    device_system(dev, hetero(24))
    dev.mpnn_init(0)
    assert dev.mpnn_harvest(1, 20, 5) == 4
    loss = dev.mpnn_loss_grad()
    assert np.isfinite(loss) and loss > 0
    with pytest.raises(cstfem_hip.CstfemError):
        dev.mpnn_harvest(0, 10, 1)

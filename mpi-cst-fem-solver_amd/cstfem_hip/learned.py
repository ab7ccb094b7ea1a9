"""Learned preconditioner: parameter layout, starting points and training on the device.

The network lives in the HIP library (csrc/mpnn.hip); this module only arranges its FP32
parameter vector and drives training through the C ABI.  Layout (in x out, row-major):

  W0 [16 x 64], b0 [64]                     encoder of the 16 (9 live) node features
  Wl [3][128 x 64], bl [3][64]              message-passing rounds on [h, mean_nb h]
  Wd [64 x 4], bd [4], Wo [64 x 4], bo [4]  decoders (diagonal / off-diagonal blocks) of
                                            [a, b, d, c]: M_ij = S_i [[a, b], [b, d]] S_j
                                            + c D_i^-1 K_ij D_j^-1,  S = D^-1/2
"""
from __future__ import annotations

import numpy as np

H, F, LIVE, ROUNDS, DEC = 64, 16, 9, 3, 4
W0 = 0
B0 = W0 + F * H
WL = B0 + H
BL = WL + ROUNDS * 2 * H * H
WD = BL + ROUNDS * H
BD = WD + H * DEC
WO = BD + DEC
BO = WO + H * DEC
NPARAMS = BO + DEC

# Degree-1 polynomial preconditioner D^-1 (I - beta K D^-1) for plane CST elasticity: beta
# slightly below 1 / lambda_max(D^-1 K) (lambda_max ~ 3.6 on the jittered plates) keeps M SPD.
POLY_BETA = 0.225


def polynomial_start(theta: np.ndarray, beta: float = POLY_BETA) -> np.ndarray:
    """Returns theta with the decoders set to the polynomial preconditioner (weights 0,
    biases (1, 0, 1, -beta) and (0, 0, 0, -beta)); the hidden layers are kept."""
    th = np.array(theta, dtype=np.float32, copy=True)
    th[WD:BD] = 0.0
    th[WO:BO] = 0.0
    th[BD:WO] = [1.0, 0.0, 1.0, -beta]
    th[BO:NPARAMS] = [0.0, 0.0, 0.0, -beta]
    return th


def white_pairs(dev, count: int, seed: int):
    """(r_k, e_k) with e_k ~ N(0, I) and r_k = K e_k (device SpMV of the constrained K)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    E = rng.standard_normal((count, dev.n))
    R = np.stack([dev.spmv(e) for e in E])
    return R.astype(np.float32), E.astype(np.float32)


def train(dev, problem, steps: int = 300, lr: float = 1e-3, pairs: int = 32, seed: int = 0,
          theta0: np.ndarray | None = None, log=None):
    """Trains on one assembled and constrained plate (problem: a pipeline.Problem).

    Loss (device): (1/K) sum_k |M r_k - e_k|_K^2 / |e_k|_K^2 over white-noise pairs; Adam.
    Returns (theta, losses)."""
    from . import pipeline
    pipeline.load_problem(dev, problem)
    pipeline.assemble_and_constrain(dev, problem)
    dev.setup()
    dev.mpnn_init(seed)
    dev.mpnn_set_params(polynomial_start(dev.mpnn_get_params()) if theta0 is None else theta0)
    R, E = white_pairs(dev, pairs, seed + 1)
    dev.mpnn_set_pairs(R, E)
    losses = []
    for it in range(steps):
        losses.append(dev.mpnn_train_step(lr))
        if log is not None and (it % 100 == 0 or it == steps - 1):
            log(f"train step {it}: loss {losses[-1]:.5f}")
    return dev.mpnn_get_params(), losses


def training_plate(n: int = 64, seed: int = 5, e_seed: int = 101):
    """A small member of the heterogeneous plate family (ln E ~ N(0,1) per element,
    plane strain, jitter 0.2h, body force) used for training."""
    from . import pipeline
    from . import PLANE_STRAIN
    return pipeline.plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=seed, E=1.0, nu=0.3,
                                  plane=PLANE_STRAIN, tx=1.0, E_sigma=1.0, E_seed=e_seed,
                                  body=(0.0, -1.0), name=f"train{n}")


def training_problem(config: int, n: int = 64):
    """A small plate of the same family as benchmark config `config` (not the benchmark
    matrix itself) to train on: the network's features are local, so it carries over to the
    large plate of the family, but not across families (plane stress against plane strain,
    graded against uniform meshes)."""
    from . import pipeline
    if config == 1:
        return pipeline.config_problem(1, n / 700.0)  # Kirsch quarter plate, n x n nodes
    if config == 0:
        return pipeline.plate_problem(n, n, seed=11, name=f"train{n}")
    return training_plate(n)

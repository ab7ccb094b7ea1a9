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
          theta0: np.ndarray | None = None, log=None, lr_final: float | None = None,
          resample: int = 0):
    """Trains on one assembled and constrained plate (problem: a pipeline.Problem).

    Loss (device): (1/K) sum_k |M r_k - e_k|_K^2 / |e_k|_K^2 over white-noise pairs; Adam.
    lr_final: the step size decays from lr to lr_final along a half cosine (constant lr when
    None); resample > 0: a new set of pairs every `resample` steps (Adam's state carries on).
    Returns (theta, losses)."""
    import math
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
        if resample > 0 and it > 0 and it % resample == 0:
            R, E = white_pairs(dev, pairs, seed + 1 + it // resample)
            dev.mpnn_set_pairs(R, E)
        step_lr = lr if lr_final is None else (
            lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / steps)))
        losses.append(dev.mpnn_train_step(step_lr))
        if log is not None and (it % 100 == 0 or it == steps - 1):
            log(f"train step {it}: loss {losses[-1]:.5f}")
    return dev.mpnn_get_params(), losses


def train_harvested(dev, plates: int = 2000, chunk: int = 250, cells: int = 128,
                    k_first: int = 1, k_last: int = 50, steps_per_chunk: int = 40,
                    lr: float = 1e-3, seed: int = 0, project: bool = True, white: int = 0,
                    theta0=None, log=None):
    """Trains on residual/error pairs that plain Jacobi-CG generates (the protocol of the
    learned preconditioner's training data): `plates` plates of cells x cells
    (pipeline.harvest_batch) in batches of `chunk`; per batch, x* of every plate by the
    batched FP64 solve, then the pairs (r_k = K e_k, e_k = x* - x_k) after plain Jacobi-CG
    iterations k_first..k_last of every plate (cstfem_batch_harvest: one pair row holds
    iteration k of all the batch's plates), with project=True each pair reduced to the error
    the multilevel coarse correction leaves (cstfem_mpnn_project_pairs: the network is the
    fine smoother of that hierarchy, so this is the part it must reduce), `white` white-noise
    pairs drawn on the device added to them when > 0, and `steps_per_chunk` Adam steps on
    the device.  Adam's state carries over between batches.
    Returns (theta, losses, info)."""
    import time
    from . import pipeline
    losses, info = [], dict(gen_s=0.0, load_s=0.0, harvest_s=0.0, train_s=0.0, pairs=0)
    nchunks = (plates + chunk - 1) // chunk
    for ci in range(nchunks):
        t0 = time.perf_counter()
        bp = pipeline.harvest_batch(min(chunk, plates - ci * chunk), seed * 1000 + ci, cells)
        t1 = time.perf_counter()
        pipeline.load_batch(dev, bp)
        if ci == 0:
            dev.mpnn_init(seed)
            dev.mpnn_set_params(polynomial_start(dev.mpnn_get_params())
                                if theta0 is None else theta0)
        t2 = time.perf_counter()
        got = dev.batch_harvest(k_first, k_last)
        if project:
            dev.mg_setup(12, 500, 1, 3)  # the hierarchy alone: Chebyshev smoother, no network
            dev.mpnn_project_pairs()
            dev.setup()  # drops the hierarchy
        if white > 0:
            dev.mpnn_add_white_pairs(white, seed * 7919 + ci + 1)
        t3 = time.perf_counter()
        for it in range(steps_per_chunk):
            losses.append(dev.mpnn_train_step(lr))
        t4 = time.perf_counter()
        info["gen_s"] += t1 - t0
        info["load_s"] += t2 - t1
        info["harvest_s"] += t3 - t2
        info["train_s"] += t4 - t3
        info["pairs"] += got * bp.count
        if log is not None:
            log(f"harvest batch {ci + 1}/{nchunks}: {bp.count} plates, {dev.n} DOFs, {got} pair "
                f"rows in {t3 - t2:.2f} s; loss {losses[-steps_per_chunk]:.4f} -> "
                f"{losses[-1]:.4f} in {t4 - t3:.2f} s")
    return dev.mpnn_get_params(), losses, info


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

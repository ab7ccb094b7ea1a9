"""The assemble-and-solve sequence of the reference's serial driver on the device.

Mirrors the phases of run_serial_case (src/driver.cpp:164-193): assembly
(constitutive_matrix, build_stiffness_pattern, assemble_stiffness), boundary conditions
(tractions, then Dirichlet elimination) and the Jacobi-CG solve, and the problem
definitions of the benchmark plates.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import Device, PLANE_STRAIN, PLANE_STRESS, METHOD_CG_FP64, meshes


@dataclasses.dataclass
class Problem:
    """A plate with material, loads and Dirichlet constraints (the case-file keys of
    driver.cpp:126-133, plus a per-element modulus and a body force)."""
    mesh: meshes.Mesh
    E: float = 1.0
    nu: float = 0.3
    plane: int = PLANE_STRESS
    thickness: float = 1.0
    E_elem: np.ndarray | None = None
    tractions: list = dataclasses.field(default_factory=list)   # (group, dict(tx,ty | kirsch))
    dirichlet: list = dataclasses.field(default_factory=list)   # (group, component, value)
    body: tuple | None = None                                   # (bx, by)
    name: str = ""

    def dirichlet_dofs(self):
        """DOFs in the order add_dirichlet visits them (bc.cpp:118-135)."""
        dofs, vals = [], []
        for g, comp, v in self.dirichlet:
            nodes = self.mesh.groups[g].reshape(-1)
            dofs.append(2 * nodes + comp)
            vals.append(np.full(nodes.size, float(v)))
        if not dofs:
            return np.zeros(0, np.int32), np.zeros(0)
        return np.concatenate(dofs).astype(np.int32), np.concatenate(vals)


def plate_problem(nx=64, ny=64, Lx=1.0, Ly=1.0, jitter=0.2, seed=1, E=1.0, nu=0.3,
                  plane=PLANE_STRESS, tx=1.0, E_sigma=None, E_seed=7, body=None, name=""):
    """Rectangle with ux=0 on x=0, uy=0 on y=0 and traction (tx, 0) on x=Lx."""
    m = meshes.rect_plate(nx, ny, Lx, Ly, jitter, seed)
    Ee = meshes.lognormal_modulus(m.num_elements, E_sigma, E_seed) if E_sigma else None
    return Problem(mesh=m, E=E, nu=nu, plane=plane, E_elem=Ee,
                   tractions=[("right", dict(tx=tx, ty=0.0))],
                   dirichlet=[("left", 0, 0.0), ("bottom", 1, 0.0)], body=body, name=name)


def config_problem(index: int, scale: float = 1.0) -> Problem:
    """The benchmark plates; `scale` < 1 shrinks the node counts for quick checks.

    0: unit square 64x64 cells, jitter 0.2h, E=1, nu=0.3, plane stress, tx=1 on x=1.
    1: Kirsch quarter plate, hole a=1 in 10x10, 700 radial (ratio 1.005) x 700 angular
       nodes, E=200e3, nu=0.3, plane stress, symmetry BCs, sigma_inf=1 on x=10.
    2: rectangle 2x1, 1413x1413 nodes, jitter 0.2h, ln E ~ N(0,1) per element, nu=0.3,
       plane strain, BCs of 0 plus body force (0,-1).
    """
    if index == 0:
        n = max(4, int(round(64 * scale)))
        return plate_problem(n, n, seed=1, name="plate64")
    if index == 1:
        n = max(8, int(round(700 * scale)))
        m = meshes.kirsch_quarter(n, n, ratio=1.005 ** (700 / n), a=1.0, L=10.0)
        return Problem(mesh=m, E=200e3, nu=0.3, plane=PLANE_STRESS, thickness=1.0,
                       tractions=[("right", dict(tx=1.0, ty=0.0))],
                       dirichlet=[("left_sym", 0, 0.0), ("bottom_sym", 1, 0.0)],
                       name="kirsch700")
    if index == 2:
        n = max(8, int(round(1412 * scale)))
        return plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=2, E=1.0, nu=0.3,
                             plane=PLANE_STRAIN, tx=1.0, E_sigma=1.0, E_seed=22,
                             body=(0.0, -1.0), name="hetero4M")
    raise ValueError(f"no config {index}")


@dataclasses.dataclass
class BatchProblem:
    """Independent plates concatenated into one mesh (plate k owns nodes
    node_off[k]..node_off[k+1]-1 and its own E, nu and traction)."""
    mesh: meshes.Mesh
    node_off: np.ndarray
    E_elem: np.ndarray
    nu_elem: np.ndarray
    plates: list          # per-plate Problem (local numbering)
    plane: int = PLANE_STRESS
    thickness: float = 1.0

    @property
    def count(self):
        return len(self.plates)


def batch_problem(count=512, seed=3, nmin=16, nmax=160) -> BatchProblem:
    """The batched workload: plates from meshes.batch_plates, ux=0 on x=0, uy=0 on y=0,
    traction (t, 0) on x = max."""
    plates = []
    xy, tri, E, nu = [], [], [], []
    node_off = [0]
    for m, prm in meshes.batch_plates(count, seed, nmin, nmax):
        p = Problem(mesh=m, E=prm["E"], nu=prm["nu"], plane=PLANE_STRESS, thickness=prm["t"],
                    tractions=[("right", dict(tx=prm["tx"], ty=0.0))],
                    dirichlet=[("left", 0, 0.0), ("bottom", 1, 0.0)], name="plate")
        plates.append(p)
        xy.append(m.xy)
        tri.append(m.tri + node_off[-1])
        E.append(np.full(m.num_elements, prm["E"]))
        nu.append(np.full(m.num_elements, prm["nu"]))
        node_off.append(node_off[-1] + m.num_nodes)
    big = meshes.Mesh(xy=np.concatenate(xy), tri=np.concatenate(tri).astype(np.int32), groups={},
                      meta=dict(kind="batch", count=count, seed=seed))
    return BatchProblem(mesh=big, node_off=np.array(node_off, np.int32), E_elem=np.concatenate(E),
                        nu_elem=np.concatenate(nu), plates=plates)


def harvest_batch(count=250, seed=0, cells=128, Lx=2.0, Ly=1.0) -> BatchProblem:
    """Plates of the learned preconditioner's training protocol: `count` rectangles Lx x Ly
    of cells x cells (jitter 0.2h), per-element ln E ~ N(0, 1), nu = 0.3, plane strain,
    ux = 0 on x = 0, uy = 0 on y = 0, a unit traction of uniformly random direction on
    x = Lx; the seed fixes meshes, moduli and directions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    plates, xy, tri, E, nu = [], [], [], [], []
    node_off = [0]
    for k in range(count):
        s_mesh, s_e = (int(v) for v in rng.integers(0, 2 ** 31, 2))
        th = float(rng.uniform(0.0, 2.0 * np.pi))
        m = meshes.rect_plate(cells, cells, Lx, Ly, 0.2, s_mesh)
        Ee = meshes.lognormal_modulus(m.num_elements, 1.0, s_e)
        p = Problem(mesh=m, E=1.0, nu=0.3, plane=PLANE_STRAIN, thickness=1.0, E_elem=Ee,
                    tractions=[("right", dict(tx=float(np.cos(th)), ty=float(np.sin(th))))],
                    dirichlet=[("left", 0, 0.0), ("bottom", 1, 0.0)], name="train")
        plates.append(p)
        xy.append(m.xy)
        tri.append(m.tri + node_off[-1])
        E.append(Ee)
        nu.append(np.full(m.num_elements, 0.3))
        node_off.append(node_off[-1] + m.num_nodes)
    big = meshes.Mesh(xy=np.concatenate(xy), tri=np.concatenate(tri).astype(np.int32), groups={},
                      meta=dict(kind="harvest", count=count, seed=seed, cells=cells))
    return BatchProblem(mesh=big, node_off=np.array(node_off, np.int32), E_elem=np.concatenate(E),
                        nu_elem=np.concatenate(nu), plates=plates, plane=PLANE_STRAIN)


def load_batch(dev: Device, bp: BatchProblem, times: "PhaseTimes | None" = None):
    """Uploads the concatenated plates and assembles/constrains them on the device."""
    t0 = time.perf_counter()
    dev.upload_mesh(bp.mesh)
    dev.set_material_fields(bp.E_elem, bp.nu_elem, bp.plane, bp.thickness)
    t1 = time.perf_counter()
    dev.build_stiffness_pattern()
    dev.batch_define(bp.node_off)
    t2 = time.perf_counter()
    dev.assemble_stiffness()
    t3 = time.perf_counter()
    dev.set_rhs(None)  # f = 0 on the device
    dofs, vals = [], []
    for k, p in enumerate(bp.plates):
        off = int(bp.node_off[k])
        for g, kw in p.tractions:
            dev.add_traction(p.mesh.groups[g] + off, **kw)
        d, v = p.dirichlet_dofs()
        dofs.append(d + 2 * off)
        vals.append(v)
    dev.apply_dirichlet(np.concatenate(dofs), np.concatenate(vals))
    dev.setup()
    t4 = time.perf_counter()
    if times is not None:
        times.upload += t1 - t0
        times.pattern += t2 - t1
        times.assembly += t3 - t2
        times.bcs += t4 - t3


@dataclasses.dataclass
class PhaseTimes:
    upload: float = 0.0
    pattern: float = 0.0
    assembly: float = 0.0
    bcs: float = 0.0
    solve: float = 0.0


def load_problem(dev: Device, prob: Problem):
    """Copies mesh and material into HBM (the reference's `read` phase)."""
    dev.upload_mesh(prob.mesh)
    dev.set_material(prob.E, prob.nu, prob.plane, prob.thickness, prob.E_elem)


def assemble_and_constrain(dev: Device, prob: Problem, times: PhaseTimes | None = None,
                           multilevel: bool = False):
    """Pattern, assembly, loads and Dirichlet elimination, all on the device.  With
    multilevel, the fine level's aggregation starts right after the pattern (mg_prepare) and
    runs on a host thread meanwhile."""
    t0 = time.perf_counter()
    dev.build_stiffness_pattern()
    if multilevel:
        dev.mg_prepare()
    t1 = time.perf_counter()
    dev.assemble_stiffness()
    t2 = time.perf_counter()
    dev.set_rhs(None)  # f = 0 on the device
    for g, kw in prob.tractions:
        dev.add_traction(prob.mesh.groups[g], **kw)
    if prob.body is not None:
        dev.add_body_force(*prob.body)
    d, v = prob.dirichlet_dofs()
    dev.apply_dirichlet(d, v)
    t3 = time.perf_counter()
    if times is not None:
        times.pattern += t1 - t0
        times.assembly += t2 - t1
        times.bcs += t3 - t2


def solve(dev: Device, rtol=1e-10, max_iter=20000, method=METHOD_CG_FP64, times=None, **kw):
    t0 = time.perf_counter()
    dev.setup()
    x, st = dev.solve(rtol=rtol, max_iter=max_iter, method=method, **kw)
    if times is not None:
        times.solve += time.perf_counter() - t0
    return x, st

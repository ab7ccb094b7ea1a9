"""Seeded synthetic meshes for the benchmark plates.

The reference reads Gmsh meshes (src/mesh_io.cpp:114-356); the workloads here are structured
plates generated in memory with the same data model as the reference's ``Mesh``
(include/cstfem/mesh.hpp:20-31): node coordinates ``xy`` (2 per node, file order = index),
counter-clockwise triangles (3 per element) and named boundary line groups (edge lists).

Generators:
  * :func:`rect_plate` -- rectangle [0,Lx]x[0,Ly] with (nx+1)x(ny+1) nodes, every cell split
    along the same diagonal, interior nodes jittered by U(-j*h, j*h) per coordinate.
    Groups ``left`` (x=0), ``bottom`` (y=0), ``right`` (x=Lx), ``top`` (y=Ly).
  * :func:`kirsch_quarter` -- quarter plate with a hole of radius ``a`` in an LxL square,
    structured radial x angular nodes mapped from the annulus to the square, radial geometric
    grading.  Groups ``hole``, ``left_sym``, ``bottom_sym``, ``right``, ``top`` (the names of
    the reference's Kirsch meshes, cases/kirsch_L0.case).
  * :func:`batch_plates` -- the 512 uneven plates of the batched workload.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Mesh:
    xy: np.ndarray            # (nn, 2) float64
    tri: np.ndarray           # (ne, 3) int32, counter-clockwise
    groups: dict              # name -> (k, 2) int32 edges
    E_elem: np.ndarray | None = None   # (ne,) float64 per-element Young's modulus
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def num_nodes(self) -> int:
        return int(self.xy.shape[0])

    @property
    def num_elements(self) -> int:
        return int(self.tri.shape[0])

    @property
    def num_dofs(self) -> int:
        return 2 * self.num_nodes

    def group_nodes(self, name: str) -> np.ndarray:
        """Nodes of a line group in the order the reference visits them
        (every node of every edge, bc.cpp:118-122)."""
        return self.groups[name].reshape(-1)


def _grid_tris(nx: int, ny: int) -> np.ndarray:
    """Cell (i, j) with corners a=(i,j), b=(i+1,j), c=(i+1,j+1), d=(i,j+1) -> (a,b,c), (a,c,d)."""
    i = np.arange(nx, dtype=np.int64)
    j = np.arange(ny, dtype=np.int64)
    I, J = np.meshgrid(i, j)  # J rows, I columns: cell order is row-major, x fastest
    a = (J * (nx + 1) + I).reshape(-1)
    b = a + 1
    d = a + (nx + 1)
    c = d + 1
    tri = np.empty((2 * a.size, 3), dtype=np.int32)
    tri[0::2] = np.stack([a, b, c], axis=1)
    tri[1::2] = np.stack([a, c, d], axis=1)
    return tri


def rect_plate(nx: int, ny: int, Lx: float = 1.0, Ly: float = 1.0, jitter: float = 0.2,
               seed: int = 0) -> Mesh:
    """Structured rectangle with nodes n = j*(nx+1) + i; interior nodes jittered."""
    hx, hy = Lx / nx, Ly / ny
    xs = np.arange(nx + 1) * hx
    ys = np.arange(ny + 1) * hy
    xs[-1], ys[-1] = Lx, Ly
    X, Y = np.meshgrid(xs, ys)
    xy = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1).astype(np.float64)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        d = rng.uniform(-1.0, 1.0, size=xy.shape)
        ii = np.tile(np.arange(nx + 1), ny + 1)
        jj = np.repeat(np.arange(ny + 1), nx + 1)
        interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)
        xy[interior, 0] += jitter * hx * d[interior, 0]
        xy[interior, 1] += jitter * hy * d[interior, 1]
    tri = _grid_tris(nx, ny)
    W = nx + 1
    bottom = np.arange(nx, dtype=np.int32)
    top = (ny * W + np.arange(nx)).astype(np.int32)
    left = (np.arange(ny) * W).astype(np.int32)
    right = (np.arange(ny) * W + nx).astype(np.int32)
    groups = {
        "bottom": np.stack([bottom, bottom + 1], axis=1),
        "top": np.stack([top, top + 1], axis=1),
        "left": np.stack([left, left + W], axis=1),
        "right": np.stack([right, right + W], axis=1),
    }
    groups = {k: np.ascontiguousarray(v, dtype=np.int32) for k, v in groups.items()}
    return Mesh(xy=xy, tri=tri, groups=groups,
                meta=dict(kind="rect", nx=nx, ny=ny, Lx=Lx, Ly=Ly, jitter=jitter, seed=seed))


def kirsch_quarter(n_rad: int = 700, n_ang: int = 700, ratio: float = 1.005, a: float = 1.0,
                   L: float = 10.0) -> Mesh:
    """Quarter plate with a hole, radial (i) x angular (j) structured nodes, n = j*n_rad + i.

    The outer boundary is walked from (L, 0) up the right edge to the corner (L, L) (node
    j = n_ang//2 - 1... placed exactly at the corner) and along the top edge to (0, L); every
    angular line joins the hole point at the outer point's polar angle to that outer point,
    with radial parameter s_i = (ratio^i - 1)/(ratio^(n_rad-1) - 1).
    """
    jc = n_ang // 2 - 1 if n_ang % 2 == 0 else n_ang // 2  # corner index
    j = np.arange(n_ang, dtype=np.float64)
    X = np.where(j <= jc, L, L * (n_ang - 1 - j) / (n_ang - 1 - jc))
    Y = np.where(j <= jc, L * j / jc, L)
    X[jc], Y[jc] = L, L
    X[-1] = 0.0
    theta = np.arctan2(Y, X)
    theta[0], theta[-1] = 0.0, np.pi / 2
    hx, hy = a * np.cos(theta), a * np.sin(theta)
    hx[-1], hy[0] = 0.0, 0.0
    i = np.arange(n_rad, dtype=np.float64)
    s = (ratio ** i - 1.0) / (ratio ** (n_rad - 1) - 1.0) if ratio != 1.0 else i / (n_rad - 1)
    s[-1] = 1.0
    S, _ = np.meshgrid(s, j)  # rows j, columns i
    x = hx[:, None] + S * (X - hx)[:, None]
    y = hy[:, None] + S * (Y - hy)[:, None]
    x[-1, :] = 0.0  # left symmetry line exactly x = 0
    y[0, :] = 0.0   # bottom symmetry line exactly y = 0
    xy = np.stack([x.reshape(-1), y.reshape(-1)], axis=1)
    tri = _grid_tris(n_rad - 1, n_ang - 1)  # radial = "x", angular = "y": CCW since e_r x e_t = +z
    W = n_rad
    jr = np.arange(n_ang - 1)
    ir = np.arange(n_rad - 1)
    hole = np.stack([jr * W, (jr + 1) * W], axis=1)
    bottom = np.stack([ir, ir + 1], axis=1)
    left = np.stack([(n_ang - 1) * W + ir, (n_ang - 1) * W + ir + 1], axis=1)
    jright = np.arange(jc)
    right = np.stack([jright * W + W - 1, (jright + 1) * W + W - 1], axis=1)
    jtop = np.arange(jc, n_ang - 1)
    top = np.stack([jtop * W + W - 1, (jtop + 1) * W + W - 1], axis=1)
    groups = {"hole": hole, "bottom_sym": bottom, "left_sym": left, "right": right, "top": top}
    groups = {k: np.ascontiguousarray(v, dtype=np.int32) for k, v in groups.items()}
    return Mesh(xy=xy, tri=tri, groups=groups,
                meta=dict(kind="kirsch", n_rad=n_rad, n_ang=n_ang, ratio=ratio, a=a, L=L))


def lognormal_modulus(ne: int, sigma: float, seed: int) -> np.ndarray:
    """Per-element E with ln E ~ N(0, sigma^2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(sigma * rng.standard_normal(ne))


def batch_plates(count: int = 512, seed: int = 3, nmin: int = 16, nmax: int = 160):
    """The batched workload's plates: nx, ny ~ U{nmin..nmax} cells, height 1, width nx/ny
    (square cells on average), jitter 0.15h, E ~ logU[0.1, 10], nu ~ U[0.2, 0.45],
    traction magnitude ~ U[0.5, 2] on x = max.  Returns a list of (Mesh, params dict)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        nx = int(rng.integers(nmin, nmax + 1))
        ny = int(rng.integers(nmin, nmax + 1))
        E = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
        nu = float(rng.uniform(0.2, 0.45))
        tx = float(rng.uniform(0.5, 2.0))
        m = rect_plate(nx, ny, Lx=nx / ny, Ly=1.0, jitter=0.15, seed=seed * 100003 + k)
        out.append((m, dict(E=E, nu=nu, tx=tx, plane_strain=0, t=1.0)))
    return out

"""GPU parity of the element stiffness entry, the device tractions and the Dirichlet
elimination that visits only the constrained rows and their neighbours.

Bars: bit-exact against the oracle (the reference's operation order); the Ke KAT of
tests/test_element.cpp:60-84 to 1e-14; rigid modes and symmetry as tests/test_element.cpp:86-119.
"""
import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import meshes, pipeline

pytestmark = pytest.mark.gpu


# ---- element stiffness (element.cpp:37-55)

def test_ke_reference_triangle_kat(dev):
    """test_element.cpp:60-84: (0,0),(1,0),(0,1), E=1, nu=0, t=1, plane stress."""
    expected = np.array([[0.75, 0.25, -0.5, -0.25, -0.25, 0.0],
                         [0.25, 0.75, 0.0, -0.25, -0.25, -0.5],
                         [-0.5, 0.0, 0.5, 0.0, 0.0, 0.0],
                         [-0.25, -0.25, 0.0, 0.25, 0.25, 0.0],
                         [-0.25, -0.25, 0.0, 0.25, 0.25, 0.0],
                         [0.0, -0.5, 0.0, 0.0, 0.0, 0.5]])
    D = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.5])  # constitutive(1, 0, stress)
    K = dev.element_stiffness(np.array([[0, 0, 1, 0, 0, 1]], float), D, 1.0)[0]
    assert np.max(np.abs(K - expected)) <= 1e-14


def random_triangles(rng, count):
    """CCW triangles of varied shape and size around random centres."""
    c = rng.uniform(-5, 5, (count, 1, 2))
    r = 10.0 ** rng.uniform(-2, 1, (count, 1, 1))
    ang = np.sort(rng.uniform(0, 2 * np.pi, (count, 3)), axis=1)
    pts = c + r * np.stack([np.cos(ang), np.sin(ang)], axis=2)
    return pts.reshape(count, 6)


@pytest.mark.parametrize("strain", [0, 1])
def test_ke_bitwise_and_properties(dev, oracle, strain):
    rng = np.random.default_rng(12345 + strain)
    xy = random_triangles(rng, 2000)
    D = oracle.constitutive(210e9, 0.3, strain)
    K = dev.element_stiffness(xy, D, 0.01)
    for e in range(0, 2000, 7):
        assert np.array_equal(K[e], oracle.stiffness(xy[e], D, 0.01)), e
    kmax = np.max(np.abs(K), axis=(1, 2))
    assert np.all(np.abs(K - K.transpose(0, 2, 1)) <= 1e-12 * kmax[:, None, None])
    x, y = xy[:, 0::2], xy[:, 1::2]
    tx = np.zeros((2000, 6))
    tx[:, 0::2] = 1.0
    ty = np.zeros((2000, 6))
    ty[:, 1::2] = 1.0
    rot = np.zeros((2000, 6))
    rot[:, 0::2], rot[:, 1::2] = -y, x
    rmax = np.max(np.abs(xy), axis=1)
    for mode, scale in ((tx, kmax), (ty, kmax), (rot, kmax * rmax)):
        r = np.einsum("eij,ej->ei", K, mode)
        assert np.all(np.abs(r) <= 1e-12 * scale[:, None])
    # a constant non-zero strain has positive energy
    u = np.zeros((2000, 6))
    u[:, 0::2] = 1e-3 * x + 2e-4 * y
    u[:, 1::2] = -1e-4 * x + 5e-4 * y
    assert np.all(np.einsum("ei,eij,ej->e", u, K, u) > 0.0)


def test_ke_empty_batch(dev):
    assert dev.element_stiffness(np.zeros((0, 6)), np.eye(3), 1.0).shape == (0, 6, 6)


# ---- tractions on the device (bc.cpp:190-253)

def fresh_rhs(dev, prob):
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    dev.assemble_stiffness()
    dev.set_rhs(None)


@pytest.mark.parametrize("nx,ny", [(200, 150), (37, 300)])
def test_traction_all_sides_bitwise(dev, oracle, nx, ny):
    """Every side of a jittered plate, traction components of both signs, reversed edge
    orientation on two sides, accumulated onto a non-zero f."""
    prob = pipeline.plate_problem(nx, ny, Lx=1.7, Ly=0.9, jitter=0.2, seed=nx)
    m = prob.mesh
    fresh_rhs(dev, prob)
    rng = np.random.default_rng(nx * ny)
    f0 = rng.standard_normal(m.num_dofs)
    dev.set_rhs(f0)
    f = f0.copy()
    loads = [("right", 1.5, -0.25), ("top", -0.75, 2.0), ("left", 0.3, 0.1), ("bottom", 0.0, -1.0)]
    for k, (g, tx, ty) in enumerate(loads):
        edges = m.groups[g]
        if k % 2:
            edges = edges[:, ::-1]
        dev.add_traction(edges, tx=tx, ty=ty)
        oracle.add_traction(m.xy, m.tri, edges, f, t=prob.thickness, tx=tx, ty=ty)
    assert np.array_equal(dev.download_rhs(), f)


def test_traction_kirsch_generator_bitwise(dev, oracle):
    """Kirsch traction (host function at device Gauss points) on a generated quarter plate."""
    m = meshes.kirsch_quarter(60, 60, ratio=1.03, a=1.0, L=10.0)
    prob = pipeline.Problem(mesh=m, E=200e3, nu=0.3, dirichlet=[("left_sym", 0, 0.0)])
    fresh_rhs(dev, prob)
    f = np.zeros(m.num_dofs)
    for g in ("right", "top"):
        dev.add_traction(m.groups[g], kirsch=(1.0, 1.0))
        oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=1.0, kirsch=(1.0, 1.0))
    assert np.array_equal(dev.download_rhs(), f)


def test_traction_function_form(dev, oracle):
    """The TractionFn form: a constant function equals the constant form bit for bit, and a
    linear field's nodal forces sum to its exact resultant (2-point Gauss is exact)."""
    prob = pipeline.plate_problem(64, 48, Lx=2.0, jitter=0.2, seed=3)
    m = prob.mesh
    fresh_rhs(dev, prob)
    dev.add_traction(m.groups["top"], tx=0.7, ty=-1.3)
    f_const = dev.download_rhs()
    dev.set_rhs(None)
    seen = []

    def const(x, y, nx, ny):
        seen.append((x.size, float(np.max(np.abs(nx * nx + ny * ny - 1.0)))))
        return np.full_like(x, 0.7), np.full_like(x, -1.3)

    dev.add_traction_fn(m.groups["top"], const)
    assert np.array_equal(dev.download_rhs(), f_const)
    assert seen[0][0] == 2 * m.groups["top"].shape[0] and seen[0][1] < 1e-15
    dev.set_rhs(None)
    dev.add_traction_fn(m.groups["top"], lambda x, y, nx, ny: (2.0 * x + 1.0, ny * y))
    f = dev.download_rhs()
    # resultant of t = (2x + 1, ny y) over each straight edge: len * t(midpoint)
    e = m.groups["top"]
    xa, ya, xb, yb = m.xy[e[:, 0], 0], m.xy[e[:, 0], 1], m.xy[e[:, 1], 0], m.xy[e[:, 1], 1]
    ln = np.hypot(xb - xa, yb - ya)
    ny_out = np.abs(xb - xa) / ln  # top side: the outward normal points up
    Fx = np.sum(ln * (2.0 * 0.5 * (xa + xb) + 1.0))
    Fy = np.sum(ln * ny_out * 0.5 * (ya + yb))
    assert abs(f[0::2].sum() - Fx) <= 1e-12 * abs(Fx)
    assert abs(f[1::2].sum() - Fy) <= 1e-12 * abs(Fy)


def test_traction_bad_edge_leaves_f(dev):
    prob = pipeline.plate_problem(20, 20, seed=1)
    m = prob.mesh
    fresh_rhs(dev, prob)
    f0 = np.arange(m.num_dofs, dtype=float)
    dev.set_rhs(f0)
    interior = np.array([[21 * 5 + 5, 21 * 5 + 6]], np.int32)  # two triangles share it
    edges = np.concatenate([m.groups["right"], interior])
    with pytest.raises(cstfem_hip.CstfemError, match="is adjacent to 2 triangles"):
        dev.add_traction(edges, tx=1.0)
    with pytest.raises(cstfem_hip.CstfemError, match="is adjacent to 0 triangles"):
        dev.add_traction(np.array([[0, 440]], np.int32), tx=1.0)
    assert np.array_equal(dev.download_rhs(), f0)


def test_traction_repeated_group_bitwise(dev, oracle):
    """A constant traction on a group applied before on the same mesh reuses its checked index
    lists (no host packing, upload or wait): the same f bit for bit, other values and groups
    too; a bad group still raises after good ones; a new mesh starts afresh."""
    prob = pipeline.plate_problem(90, 70, Lx=1.3, jitter=0.2, seed=5)
    m = prob.mesh
    fresh_rhs(dev, prob)
    for rep in range(3):
        dev.set_rhs(None)
        f = np.zeros(m.num_dofs)
        for g, tx, ty in (("right", 1.0 + rep, -0.5), ("top", -0.25, 2.0 - rep)):
            dev.add_traction(m.groups[g], tx=tx, ty=ty)
            oracle.add_traction(m.xy, m.tri, m.groups[g], f, t=prob.thickness, tx=tx, ty=ty)
        assert np.array_equal(dev.download_rhs(), f), rep
    # the reversed orientation is another group (its own check and lists)
    dev.set_rhs(None)
    f = np.zeros(m.num_dofs)
    dev.add_traction(m.groups["right"][:, ::-1], tx=0.5, ty=0.5)
    oracle.add_traction(m.xy, m.tri, m.groups["right"][:, ::-1], f, t=prob.thickness, tx=0.5,
                        ty=0.5)
    assert np.array_equal(dev.download_rhs(), f)
    with pytest.raises(cstfem_hip.CstfemError, match="is adjacent to 0 triangles"):
        dev.add_traction(np.array([[0, 90 * 71 - 1]], np.int32), tx=1.0)
    with pytest.raises(cstfem_hip.CstfemError, match="is adjacent to 0 triangles"):
        dev.add_traction(np.array([[0, 90 * 71 - 1]], np.int32), tx=1.0)
    # another mesh: the same node numbers are other edges, checked again
    prob2 = pipeline.plate_problem(40, 30, seed=2)
    fresh_rhs(dev, prob2)
    f = np.zeros(prob2.mesh.num_dofs)
    dev.add_traction(prob2.mesh.groups["right"], tx=1.0, ty=0.0)
    oracle.add_traction(prob2.mesh.xy, prob2.mesh.tri, prob2.mesh.groups["right"], f,
                        t=prob2.thickness, tx=1.0, ty=0.0)
    assert np.array_equal(dev.download_rhs(), f)


# ---- Dirichlet elimination (bc.cpp:137-166)

def test_dirichlet_repeated_calls_vs_oracle(dev, oracle):
    """Two eliminations in a row (prescribed non-zero values, interior nodes included) equal
    the reference's two apply_dirichlet calls bit for bit."""
    prob = pipeline.plate_problem(90, 70, jitter=0.2, seed=8, E_sigma=1.0)
    m = prob.mesh
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    rp, c = dev.download_pattern()
    dev.assemble_stiffness()
    v = dev.download_values()
    rng = np.random.default_rng(5)
    f = rng.standard_normal(m.num_dofs)
    dev.set_rhs(f)
    d1 = np.concatenate([2 * np.unique(m.groups["left"]),
                         2 * np.unique(m.groups["bottom"]) + 1]).astype(np.int32)
    v1 = rng.uniform(-1, 1, d1.size)
    d2 = rng.choice(m.num_dofs, 300, replace=False).astype(np.int32)
    d2 = d2[~np.isin(d2, d1)]
    v2 = rng.uniform(-2, 2, d2.size)
    dev.apply_dirichlet(d1, v1)
    dev.apply_dirichlet(d2, v2)
    K1, b1 = oracle.apply_dirichlet(rp, c, v, f, d1, v1)
    K2, b2 = oracle.apply_dirichlet(rp, c, K1, b1, d2, v2)
    assert np.array_equal(dev.download_values(), K2)
    assert np.array_equal(dev.download_rhs(), b2)


def test_dirichlet_nonsymmetric_pattern_vs_oracle(dev, oracle):
    """A generic CSR upload whose pattern is not symmetric: every row is visited."""
    rng = np.random.default_rng(9)
    n = 301
    A = np.where(rng.uniform(size=(n, n)) < 0.03, rng.standard_normal((n, n)), 0.0)
    A += np.diag(rng.uniform(1, 2, n) * n)
    rp, c, v = [0], [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        c.extend(nz.tolist())
        v.extend(A[i, nz].tolist())
        rp.append(len(c))
    rp, c, v = np.array(rp, np.int32), np.array(c, np.int32), np.array(v)
    dev.upload_csr(rp, c, v)
    b = rng.standard_normal(n)
    dev.set_rhs(b)
    d = rng.choice(n, 40, replace=False).astype(np.int32)
    u = rng.uniform(-1, 1, 40)
    dev.apply_dirichlet(d, u)
    K, bb = oracle.apply_dirichlet(rp, c, v, b, d, u)
    assert np.array_equal(dev.download_values(), K)
    assert np.array_equal(dev.download_rhs(), bb)


def test_dirichlet_duplicates_and_first_conflict(dev):
    """DirichletDofs::add: a repeated DOF with its value is accepted; the first conflict in
    input order is reported with the first value given (bc.cpp:107-116)."""
    rp = np.array([0, 2, 5, 7], np.int32)
    c = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v = np.array([4, 1, 1, 3, 1, 1, 2], float)
    dev.upload_csr(rp, c, v)
    dev.set_rhs(np.ones(3))
    dev.apply_dirichlet([1, 1], [2.0, 2.0])
    assert np.array_equal(dev.download_rhs(), [-1.0, 6.0, -1.0])
    with pytest.raises(cstfem_hip.CstfemError,
                       match=r"node 1 \(0-based\) component ux: 5 vs 6"):
        dev.apply_dirichlet([0, 2, 1, 2, 0], [1.0, 5.0, 0.0, 6.0, 7.0])

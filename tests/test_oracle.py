"""Pins the oracle (oracle/cst_oracle.c) before it is trusted as the checker.

Each oracle function is checked against what the reference's own tests hold (hand KATs,
closed forms, Cholesky, golden fixtures written from the reference's compiled core) and, in
this container, bitwise against the reference core itself (oracle/_ref).  The checks are
chosen so that a dropped term, a wrong sign or a transposed index fails one of them.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from cstfem_hip import meshes, pipeline

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden():
    with open(os.path.join(GOLD, "golden.json")) as fh:
        return json.load(fh)


def kirsch_mesh(lvl):
    z = np.load(os.path.join(GOLD, f"kirsch_L{lvl}_mesh.npz"))
    return meshes.Mesh(xy=z["xy"], tri=z["tri"],
                       groups={k[6:]: z[k] for k in z.files if k.startswith("group_")})


# ---- element (test_element.cpp) -------------------------------------------------------

def test_ke_reference_triangle(oracle):
    """test_element.cpp:60-84: Ke of (0,0),(1,0),(0,1), E=1, nu=0, t=1 by hand."""
    expected = np.array([[0.75, 0.25, -0.5, -0.25, -0.25, 0.0],
                         [0.25, 0.75, 0.0, -0.25, -0.25, -0.5],
                         [-0.5, 0.0, 0.5, 0.0, 0.0, 0.0],
                         [-0.25, -0.25, 0.0, 0.25, 0.25, 0.0],
                         [-0.25, -0.25, 0.0, 0.25, 0.25, 0.0],
                         [0.0, -0.5, 0.0, 0.0, 0.0, 0.5]])
    K = oracle.stiffness([0, 0, 1, 0, 0, 1], oracle.constitutive(1.0, 0.0, 0), 1.0)
    assert np.max(np.abs(K - expected)) <= 1e-14


def test_constitutive_closed_form(oracle):
    """test_element.cpp:121-136."""
    Ds = oracle.constitutive(1.0, 0.3, 0)
    De = oracle.constitutive(1.0, 0.3, 1)
    Ds_ref = np.array([1, 0.3, 0, 0.3, 1, 0, 0, 0, 0.35]) / 0.91
    De_ref = np.array([0.7, 0.3, 0, 0.3, 0.7, 0, 0, 0, 0.2]) / 0.52
    assert np.max(np.abs(Ds - Ds_ref)) <= 1e-15
    assert np.max(np.abs(De - De_ref)) <= 1e-15


@pytest.mark.parametrize("strain", [0, 1])
def test_ke_properties(oracle, strain):
    """test_element.cpp:86-119: symmetry, three rigid modes in the null space, positive
    energy of a constant strain field, on random counter-clockwise triangles."""
    rng = np.random.default_rng(12345)
    D = oracle.constitutive(210e9, 0.3, strain)
    done = 0
    while done < 20:
        xy = rng.uniform(-2, 3, 6)
        A = 0.5 * ((xy[2] - xy[0]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[3] - xy[1]))
        if abs(A) < 0.5:
            continue
        if A < 0:
            xy[[2, 3, 4, 5]] = xy[[4, 5, 2, 3]]
        K = oracle.stiffness(xy, D, 0.01)
        kmax = np.max(np.abs(K))
        assert np.max(np.abs(K - K.T)) <= 1e-12 * kmax
        rmax = np.max(np.abs(xy))
        tx = np.tile([1.0, 0.0], 3)
        ty = np.tile([0.0, 1.0], 3)
        rot = np.ravel(np.column_stack([-xy[1::2], xy[0::2]]))
        assert np.max(np.abs(K @ tx)) <= 1e-12 * kmax
        assert np.max(np.abs(K @ ty)) <= 1e-12 * kmax
        assert np.max(np.abs(K @ rot)) <= 1e-12 * kmax * rmax
        u = np.ravel(np.column_stack([1e-3 * xy[0::2] + 2e-4 * xy[1::2],
                                      -1e-4 * xy[0::2] + 5e-4 * xy[1::2]]))
        assert u @ K @ u > 0
        done += 1


def test_stiffness_matches_reference_bitwise(oracle, ref):
    rng = np.random.default_rng(7)
    for strain in (0, 1):
        D = oracle.constitutive(3.7, 0.27, strain)
        assert np.array_equal(D, ref.constitutive(3.7, 0.27, strain))
        for _ in range(50):
            xy = rng.uniform(-1, 1, 6)
            assert np.array_equal(oracle.stiffness(xy, D, 0.3), ref.stiffness(xy, D, 0.3))


# ---- pattern and assembly (test_csr.cpp) ----------------------------------------------

def edge_count(tri):
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0).shape[0]


@pytest.mark.parametrize("lvl", [0, 1])
def test_pattern_kirsch(oracle, lvl):
    """test_csr.cpp:36-62: nnz = 4 (nodes + 2 edges), sorted columns, diagonal present;
    and the reference's digests (golden fixtures)."""
    m = kirsch_mesh(lvl)
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    assert rp[-1] == 4 * (m.num_nodes + 2 * edge_count(m.tri))
    g = golden()[f"kirsch_L{lvl}"]
    assert rp[-1] == g["nnz"]
    assert sha(rp) == g["sha_row_ptr"] and sha(c) == g["sha_cols"]
    for i in range(rp.size - 1):
        row = c[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row


@pytest.mark.parametrize("strain", [0, 1])
def test_global_symmetry_rigid_modes(oracle, strain):
    """test_csr.cpp:64-98 on kirsch_L0: symmetry and rigid modes <= 1e-12."""
    m = kirsch_mesh(0)
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=210e9, nu=0.3, strain=strain, t=0.01)
    n = rp.size - 1
    import scipy.sparse as sp
    K = sp.csr_matrix((v, c, rp), shape=(n, n))
    assert abs(K - K.T).max() <= 1e-12 * np.max(np.abs(v))
    knorm = np.max(np.abs(K).sum(axis=1))
    xy = m.xy
    for r in (np.tile([1.0, 0.0], m.num_nodes), np.tile([0.0, 1.0], m.num_nodes),
              np.ravel(np.column_stack([-xy[:, 1], xy[:, 0]]))):
        assert np.max(np.abs(K @ r)) / (knorm * np.max(np.abs(r))) <= 1e-12


def test_assembly_golden(oracle):
    """Values, loads and eliminated system on the reference's Kirsch meshes and the
    synthetic plate equal the reference's bit for bit (digests from the compiled core)."""
    g = golden()
    for lvl in (0, 1):
        m = kirsch_mesh(lvl)
        rp, c = oracle.build_pattern(m.num_nodes, m.tri)
        v = oracle.assemble(m.xy, m.tri, rp, c, E=210e9, nu=0.3, strain=0, t=0.01)
        assert sha(v) == g[f"kirsch_L{lvl}"]["sha_vals"]
        f = np.zeros(m.num_dofs)
        for grp in ("right", "top"):
            oracle.add_traction(m.xy, m.tri, m.groups[grp], f, t=0.01, kirsch=(100e6, 0.1))
        assert sha(f) == g[f"kirsch_L{lvl}"]["sha_f"]
    prob = pipeline.config_problem(0)
    m = prob.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=1.0, nu=0.3)
    assert sha(v) == g["plate64"]["sha_vals"] and sha(c) == g["plate64"]["sha_cols"]


def test_hetero_reduces_to_uniform(oracle):
    """Per-element E equal to the scalar E reproduces the uniform assembly bit for bit;
    doubling every E_e doubles every value exactly."""
    m = meshes.rect_plate(20, 13, jitter=0.2, seed=4)
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=2.5, nu=0.3)
    v1 = oracle.assemble(m.xy, m.tri, rp, c, E=99.0, nu=0.3, E_elem=np.full(m.num_elements, 2.5))
    assert np.array_equal(v, v1)
    v2 = oracle.assemble(m.xy, m.tri, rp, c, E=99.0, nu=0.3, E_elem=np.full(m.num_elements, 5.0))
    assert np.array_equal(v2, 2 * v)


def test_two_material_bar(oracle):
    """A bar of two materials in series under end traction: the discrete solution is
    exactly piecewise linear, u(L) = sigma (L1/E1 + L2/E2) (1D closed form, nu = 0)."""
    nx, ny = 20, 4
    m = meshes.rect_plate(nx, ny, Lx=2.0, Ly=0.5, jitter=0.0)
    cx = m.xy[m.tri].mean(axis=1)[:, 0]
    E_elem = np.where(cx < 1.0, 3.0, 0.5)
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=1.0, nu=0.0, E_elem=E_elem)
    f = np.zeros(m.num_dofs)
    oracle.add_traction(m.xy, m.tri, m.groups["right"], f, t=1.0, tx=2.0)
    left = m.groups["left"].reshape(-1)
    dofs = np.concatenate([2 * left, [2 * left[0] + 1]])
    K, b = oracle.apply_dirichlet(rp, c, v, f, dofs, np.zeros(dofs.size))
    x, st = oracle.pcg_jacobi(rp, c, K, b, 1e-13, 5000)
    ux = x[0::2]
    xs = m.xy[:, 0]
    exact = np.where(xs < 1.0, 2.0 * xs / 3.0, 2.0 / 3.0 + 2.0 * (xs - 1.0) / 0.5)
    assert np.max(np.abs(ux - exact)) <= 1e-9 * np.max(np.abs(exact))


def test_body_force_resultant(oracle):
    """Sum of nodal body loads = t * area * (bx, by)."""
    m = meshes.rect_plate(17, 9, Lx=2.0, Ly=1.0, jitter=0.2, seed=3)
    f = np.zeros(m.num_dofs)
    oracle.add_body_force(m.xy, m.tri, f, 0.7, 0.3, -1.0)
    assert abs(f[0::2].sum() - 0.7 * 2.0 * 0.3) <= 1e-13
    assert abs(f[1::2].sum() + 0.7 * 2.0) <= 1e-13


# ---- Dirichlet and tractions (test_bc.cpp) --------------------------------------------

def test_symmetric_elimination_kat(oracle):
    """test_bc.cpp:63-78 (exact equality)."""
    rp = np.array([0, 2, 5, 7], np.int32)
    c = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v = np.array([4, 1, 1, 3, 1, 1, 2], float)
    K, b = oracle.apply_dirichlet(rp, c, v, np.ones(3), [1], [2.0])
    assert np.array_equal(b, [-1.0, 6.0, -1.0])
    assert np.array_equal(K, [4, 0, 0, 3, 0, 0, 2])


def test_traction_equilibrium(oracle):
    """test_bc.cpp:107-126: resultant = t * traction * group length to 1e-12."""
    m = kirsch_mesh(0)
    t, tx, ty = 0.01, 3.0e6, -2.0e6
    for g in ("hole", "left_sym", "bottom_sym", "right", "top"):
        e = m.groups[g]
        length = np.sum(np.hypot(*(m.xy[e[:, 1]] - m.xy[e[:, 0]]).T))
        f = np.zeros(m.num_dofs)
        oracle.add_traction(m.xy, m.tri, e, f, t=t, tx=tx, ty=ty)
        assert abs(f[0::2].sum() - t * tx * length) <= 1e-12 * abs(t * tx * length)
        assert abs(f[1::2].sum() - t * ty * length) <= 1e-12 * abs(t * ty * length)


def test_dirichlet_matches_reference(oracle, ref):
    prob = pipeline.plate_problem(12, 9, seed=8, E_sigma=None)
    m = prob.mesh
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c)
    rng = np.random.default_rng(0)
    f = rng.standard_normal(m.num_dofs)
    d, val = prob.dirichlet_dofs()
    val = val + rng.uniform(-1, 1, val.size) * 0  # zero prescriptions
    dofs = np.concatenate([d, [2 * 40 + 1]])
    vals = np.concatenate([val, [0.37]])
    K1, b1 = oracle.apply_dirichlet(rp, c, v, f, dofs, vals)
    K2, b2 = ref.apply_dirichlet(rp, c, v, f, dofs, vals)
    assert np.array_equal(K1, K2) and np.array_equal(b1, b2)


# ---- PCG (test_pcg.cpp) -----------------------------------------------------------------

def random_spd(n, rng):
    M = rng.uniform(-1, 1, (n, n))
    A = M.T @ M + n * np.eye(n)
    d = 10.0 ** (3.0 * (rng.uniform(-1, 1, n) + 1.0) / 2.0)
    return A * d[:, None] * d[None, :]


def dense_to_csr(A):
    rp, c, v = [0], [], []
    for i in range(A.shape[0]):
        nz = np.nonzero(A[i])[0]
        c.extend(nz.tolist())
        v.extend(A[i, nz].tolist())
        rp.append(len(c))
    return np.array(rp, np.int32), np.array(c, np.int32), np.array(v)


@pytest.mark.parametrize("n", [5, 40, 120])
def test_pcg_vs_cholesky(oracle, n):
    """test_pcg.cpp:75-102."""
    rng = np.random.default_rng(n)
    A = random_spd(n, rng)
    b = rng.uniform(-1, 1, n)
    x, st = oracle.pcg_jacobi(*dense_to_csr(A), b, 1e-12, 1000, history=True)
    x_ref = np.linalg.solve(A, b)
    assert st["converged"] and st["rel_residual"] < 1e-12
    assert np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)) < 1e-9
    assert st["history"].size == st["iterations"] + 1 and st["history"][0] == 1.0


def test_pcg_edge_cases(oracle):
    """test_pcg.cpp:104-128: zero RHS, max_iter."""
    A = random_spd(6, np.random.default_rng(1))
    x, st = oracle.pcg_jacobi(*dense_to_csr(A), np.zeros(6))
    assert st["converged"] and st["iterations"] == 0 and not np.any(x)
    A = random_spd(30, np.random.default_rng(3))
    x, st = oracle.pcg_jacobi(*dense_to_csr(A), np.ones(30), 1e-14, 2)
    assert not st["converged"] and st["iterations"] == 2


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_pcg_baseline(oracle, threads):
    """The threaded PCG (timed CPU baseline) runs the same recurrence; only its dot-product
    summation order differs, so it converges to the same solution in about as many
    iterations; with one thread it is the sequential oracle bit for bit."""
    A = random_spd(150, np.random.default_rng(11))
    b = np.random.default_rng(12).uniform(-1, 1, 150)
    x1, s1 = oracle.pcg_jacobi(*dense_to_csr(A), b, 1e-12, 2000)
    x2, s2 = oracle.pcg_jacobi_mt(*dense_to_csr(A), b, 1e-12, 2000, threads=threads)
    assert s2["converged"] and abs(s2["iterations"] - s1["iterations"]) <= 2
    assert np.max(np.abs(x2 - x1)) <= 1e-8 * np.max(np.abs(x1))
    if threads == 1:
        assert np.array_equal(x1, x2) and s1["iterations"] == s2["iterations"]
    _, s0 = oracle.pcg_jacobi_mt(*dense_to_csr(A), np.zeros(150), threads=threads)
    assert s0["converged"] and s0["zero_rhs"] and s0["iterations"] == 0


def test_pcg_matches_reference_bitwise(oracle, ref):
    A = random_spd(60, np.random.default_rng(9))
    b = np.random.default_rng(10).uniform(-1, 1, 60)
    x1, s1 = oracle.pcg_jacobi(*dense_to_csr(A), b, 1e-12, 1000, history=True)
    x2, s2 = ref.pcg_jacobi(*dense_to_csr(A), b, 1e-12, 1000, history=True)
    assert np.array_equal(x1, x2) and s1["iterations"] == s2["iterations"]
    assert np.array_equal(s1["history"], s2["history"])


@pytest.mark.parametrize("lvl", [0, 1])
def test_kirsch_end_to_end_golden(oracle, lvl):
    """Kirsch L0/L1 at rtol 1e-12: 148 / 295 iterations and the published Kirsch metrics
    (docs/verification_results.json), solution bitwise equal to the reference's."""
    g = golden()[f"kirsch_L{lvl}"]
    m = kirsch_mesh(lvl)
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    v = oracle.assemble(m.xy, m.tri, rp, c, E=210e9, nu=0.3, t=0.01)
    f = np.zeros(m.num_dofs)
    for grp in ("right", "top"):
        oracle.add_traction(m.xy, m.tri, m.groups[grp], f, t=0.01, kirsch=(100e6, 0.1))
    d = np.concatenate([2 * m.groups["left_sym"].reshape(-1),
                        2 * m.groups["bottom_sym"].reshape(-1) + 1])
    K, b = oracle.apply_dirichlet(rp, c, v, f, d, np.zeros(d.size))
    assert sha(K) == g["sha_K"] and sha(b) == g["sha_b"]
    x, st = oracle.pcg_jacobi(rp, c, K, b, 1e-12, 20000)
    assert st["iterations"] == {0: 148, 1: 295}[lvl] == g["iterations"]
    assert sha(x) == g["sha_x"]
    km = oracle.kirsch_metrics(m.xy, m.tri, x, 100e6, 0.1, 210e9, 0.3)
    for k, val in g["kirsch"].items():
        assert km[k] == pytest.approx(val, rel=1e-12, abs=0), k
    if lvl == 1:
        assert km["scf_numerical"] == pytest.approx(3.0446101947594397, rel=1e-13)


def test_patch_tests(oracle):
    """test_patch.cpp:52-124 on kirsch_L0: Dirichlet and uniaxial Neumann patch tests
    reproduce the linear field to <1e-9; a flipped normal fails (>1e-2)."""
    m = kirsch_mesh(0)
    E, nu, t = 210e9, 0.3, 0.01
    a, bb = (1e-4, 2e-3, -1e-3), (-5e-4, 1e-3, 3e-3)
    ex = lambda x, y: (a[0] + a[1] * x + a[2] * y, bb[0] + bb[1] * x + bb[2] * y)  # noqa: E731
    rp, c = oracle.build_pattern(m.num_nodes, m.tri)
    for strain in (0, 1):
        v = oracle.assemble(m.xy, m.tri, rp, c, E=E, nu=nu, strain=strain, t=t)
        nodes = np.unique(np.concatenate([m.groups[g].reshape(-1) for g in m.groups]))
        ux, uy = ex(m.xy[nodes, 0], m.xy[nodes, 1])
        dofs = np.concatenate([2 * nodes, 2 * nodes + 1])
        K, b = oracle.apply_dirichlet(rp, c, v, np.zeros(m.num_dofs), dofs,
                                      np.concatenate([ux, uy]))
        x, st = oracle.pcg_jacobi(rp, c, K, b, 1e-12, 20000)
        exu = np.ravel(np.column_stack(ex(m.xy[:, 0], m.xy[:, 1])))
        assert st["converged"] and np.max(np.abs(x - exu)) / np.max(np.abs(exu)) < 1e-9
    # Neumann, uniaxial sxx on right/top/hole (plane stress)
    s0 = 100e6
    v = oracle.assemble(m.xy, m.tri, rp, c, E=E, nu=nu, strain=0, t=t)
    for flip in (None, "right"):
        f = np.zeros(m.num_dofs)
        for g in ("right", "top", "hole"):
            e = m.groups[g]
            # traction sigma.n with sigma = diag(s0, 0): tx = s0 nx, ty = 0 -> per edge normal
            for k in range(e.shape[0]):
                fe = np.zeros(m.num_dofs)
                oracle.add_traction(m.xy, m.tri, e[k:k + 1], fe, t=t, tx=1.0, ty=0.0)
                # recover the edge's outward normal x-component from the unit traction load
                length = np.hypot(*(m.xy[e[k, 1]] - m.xy[e[k, 0]]))
                nxv = _edge_normal_x(m, e[k])
                s = -1.0 if flip == g else 1.0
                f += fe * s0 * nxv * s
        d = np.concatenate([2 * m.groups["left_sym"].reshape(-1),
                            2 * m.groups["bottom_sym"].reshape(-1) + 1])
        K, b = oracle.apply_dirichlet(rp, c, v, f, d, np.zeros(d.size))
        x, st = oracle.pcg_jacobi(rp, c, K, b, 1e-12, 20000)
        exu = np.ravel(np.column_stack([s0 / E * m.xy[:, 0], -nu * s0 / E * m.xy[:, 1]]))
        err = np.max(np.abs(x - exu)) / np.max(np.abs(exu))
        if flip is None:
            assert err < 1e-9
        else:
            assert err > 1e-2


def _edge_normal_x(m, e):
    """Outward normal x-component of a boundary edge (normal rule of bc.cpp:227-234)."""
    a, b = e
    xa, ya = m.xy[a]
    xb, yb = m.xy[b]
    L = np.hypot(xb - xa, yb - ya)
    nx, ny = (yb - ya) / L, -(xb - xa) / L
    tri = m.tri[np.any(m.tri == a, axis=1) & np.any(m.tri == b, axis=1)][0]
    cn = [n for n in tri if n != a and n != b][0]
    if nx * (m.xy[cn, 0] - xa) + ny * (m.xy[cn, 1] - ya) > 0:
        nx = -nx
    return nx

"""Multilevel preconditioner on the GPU against the CPU oracle (oracle/mg_oracle.py).

- aggregates of every level equal the oracle's (integers, exact);
- coarse matrices (Galerkin products on the device) equal the oracle's to 1e-12 relative;
- one application of the cycle equals the oracle's cycle (FP64; reductions in a different
  order) to 1e-9 relative, for hierarchies of 2 to 5 levels and K-cycle depths 1 to all;
- the headline cycle with the learned fine smoother equals the oracle's cycle given the
  device's M and stored values (FP32 copies included) to 1e-9, and its flexible CG (z'q
  from the smoothing kernel) matches the oracle's FCG in iterations and solution;
- FCG with the multilevel preconditioner reaches the FP64 true residual in a small fraction
  of Jacobi-CG's iterations, with the Chebyshev and with the learned fine smoother;
- the trained smoother needs fewer iterations than the untrained (polynomial) one and than
  Chebyshev, on a plate it was not trained on;
- a learned smoother that overshoots (lambda_max(M A) > 1.5) is replaced by Chebyshev at
  set-up; one that breaks down during the solve hands over to the same hierarchy with the
  Chebyshev smoother; either way the solve converges;
- a good trained smoother passes the set-up check;
- the fused cycle kernels give bit-identical cycles and solves to the separate kernels;
- the set-up's wavefront kernels handle aggregates of many (member, block) pairs and coarse
  rows of more than 64 blocks (oracle), and its lane-per-aggregate kernels build the same
  hierarchy, bit for bit, as the wavefront forms alone.
No reference counterpart exists for this component.
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp

import cstfem_hip
from cstfem_hip import learned, pipeline
from oracle import mg_oracle as mgo
from oracle import mpnn_oracle as mo

pytestmark = pytest.mark.gpu

FCG = cstfem_hip.METHOD_FCG_LEARNED


def hetero(n, e_seed=22, seed=2):
    return pipeline.plate_problem(n, n, Lx=2.0, Ly=1.0, jitter=0.2, seed=seed, E=1.0, nu=0.3,
                                  plane=cstfem_hip.PLANE_STRAIN, tx=1.0, E_sigma=1.0,
                                  E_seed=e_seed, body=(0.0, -1.0))


def system(dev, prob):
    pipeline.load_problem(dev, prob)
    dev.build_stiffness_pattern()
    rp, c = dev.download_pattern()
    pipeline.assemble_and_constrain(dev, prob)
    K = dev.download_values()
    dev.setup()
    d, _ = prob.dirichlet_dofs()
    fixed = np.zeros(dev.n, bool)
    fixed[d] = True
    A = sp.csr_matrix((K, c, rp), shape=(dev.n, dev.n))
    # node graph = block pattern: scalar row 2i holds columns 2j, 2j+1 of every neighbour j
    nn = dev.n // 2
    deg = (rp[1::2][:nn] - rp[0::2][:nn]) // 2
    bptr = np.concatenate([[0], np.cumsum(deg)])
    bcol = np.concatenate([c[rp[2 * i]:rp[2 * i + 1]][0::2] // 2 for i in range(nn)])
    return A, prob.mesh.xy, fixed, bptr, bcol


@pytest.mark.parametrize("n,max_levels,coarse_max,kdepth", [(16, 2, 4096, 2), (24, 3, 40, 2),
                                                            (40, 5, 30, 1), (40, 5, 30, 9)])
def test_hierarchy_and_cycle_match_oracle(dev, n, max_levels, coarse_max, kdepth):
    A, xy, fixed, bptr, bcol = system(dev, hetero(n))
    dev.mg_setup(max_levels, coarse_max, 1, kdepth)
    levels = mgo.setup(A, xy, fixed, max_levels=max_levels, coarse_max=coarse_max,
                       ptr0=bptr, col0=bcol)
    nlev = dev.mg_levels()
    assert nlev == len(levels)
    lams = []
    for l in range(nlev):
        nl, na, nnz, lam = dev.mg_level_info(l)
        assert nl == levels[l].n
        lams.append(lam)
        if l < nlev - 1:
            agg = dev.mg_download_aggregates(l)
            np.testing.assert_array_equal(agg, levels[l].agg)
        if l >= 1:
            rp, cols, vals = dev.mg_download_matrix(l)
            Ad = sp.csr_matrix((vals, cols, rp), shape=(nl, nl))
            Ao = levels[l].A
            assert abs(Ad - Ao).max() <= 1e-12 * abs(Ao).max(), l
    rng = np.random.default_rng(5)
    for _ in range(2):
        r = rng.standard_normal(dev.n)
        z = dev.mg_apply(r)
        zo = mgo.cycle(levels, 0, r, lams, kdepth=kdepth)
        assert np.linalg.norm(z - zo) <= 1e-9 * np.linalg.norm(zo)
    # the cycle is deterministic
    np.testing.assert_array_equal(dev.mg_apply(r), z)


@pytest.fixture(scope="module")
def theta40(dev):
    """A network trained briefly (40 Adam steps from the polynomial start) on a small plate of
    the family: its decoder uses all of a, b, d, c."""
    theta, _ = learned.train(dev, learned.training_plate(24), steps=40)
    return theta


def device_m(dev, prob, A):
    """The learned M as the device stores it (csrc/mpnn.hip, k_decode): FP32 values formed
    from the decoder outputs [a, b, d, c] of every stored block in the device's operation
    order, g = (s_i a) s_j + c (float)(d_i K_ij d_j) with s = (float)sqrt(D^-1)."""
    g = mo.Graph(A.indptr, A.indices, A.data, prob.mesh.tri, dev.n // 2)
    out = dev.mpnn_download(2, len(g.bcol) * mo.DEC).reshape(-1, mo.DEC)
    a, b, d, c = (out[:, u] for u in range(4))
    s = np.sqrt(g.invd).astype(np.float32)
    i, j = g.brow, g.bcol
    si, sj = (s[2 * i], s[2 * i + 1]), (s[2 * j], s[2 * j + 1])
    di, dj = (g.invd[2 * i], g.invd[2 * i + 1]), (g.invd[2 * j], g.invd[2 * j + 1])
    blk = g.blocks
    vals = []
    for (u, v), w in zip(((0, 0), (0, 1), (1, 0), (1, 1)), (a, b, b, d)):
        t = (di[u] * blk[:, u, v] * dj[v]).astype(np.float32)
        vals.append((si[u] * w * sj[v] + c * t).astype(np.float64))
    rows = np.concatenate([2 * i, 2 * i, 2 * i + 1, 2 * i + 1])
    cols = np.concatenate([2 * j, 2 * j + 1, 2 * j, 2 * j + 1])
    return sp.csr_matrix((np.concatenate(vals), (rows, cols)), shape=A.shape)


def rounded32(A):
    return sp.csr_matrix((A.data.astype(np.float32).astype(np.float64), A.indices, A.indptr),
                         shape=A.shape)


def tail_start(sizes, kdepth):
    """First level of the device's fused single-workgroup tail (csrc/mg.hip, tail_start):
    levels below the K-cycle levels, at most 4 above the coarsest, whose vectors (5 per level,
    2 on the coarsest) fit 64 KB of LDS; None when there is no tail."""
    last = len(sizes) - 1
    t = last
    while (t - 1 >= 1 and t - 1 > kdepth and last - (t - 1) <= 4
           and 8 * (2 * sizes[last] + 5 * sum(sizes[t - 1:last])) <= 64 * 1024):
        t -= 1
    return t if t < last else None


def learned_hierarchy(dev, theta, monkeypatch, coarse32, tail, n=40, kdepth=3):
    """The device's learned-smoother hierarchy on a plate and the oracle's, with the oracle's
    row sums reading the values the device's cycle reads: FP32 copies of K (level-0
    residuals), of the coarse matrices and of level 0's P when CSTFEM_MG_COARSE32 is on,
    except on the levels of the fused tail, which keep their FP64 values."""
    prob = hetero(n, e_seed=3, seed=4)
    A, xy, fixed, bptr, bcol = system(dev, prob)
    dev.mpnn_set_params(theta)
    dev.mpnn_build()
    monkeypatch.setenv("CSTFEM_MG_COARSE32", coarse32)
    monkeypatch.setenv("CSTFEM_MG_TAIL", tail)
    dev.mg_setup(6, 40, 0, kdepth)
    levels = mgo.setup(A, xy, fixed, max_levels=6, coarse_max=40, ptr0=bptr, col0=bcol)
    assert dev.mg_levels() == len(levels) >= 4
    lams = [dev.mg_level_info(l)[3] for l in range(len(levels))]
    assert 0.5 < lams[0] <= 1.5  # the set-up kept the learned smoother
    if coarse32 == "1":
        t = tail_start([L.n for L in levels], kdepth) if tail == "1" else None
        for l, L in enumerate(levels[:-1]):
            if t is None or l < t:
                L.As = rounded32(L.A)
        # level 0's restriction and prolongation read P rounded to FP32 (the Galerkin
        # products, already formed, read it in FP64)
        levels[0].P = rounded32(levels[0].P)
    return prob, A, levels, lams, device_m(dev, prob, A)


@pytest.mark.parametrize("coarse32,tail,kdepth", [("0", "0", 3), ("1", "0", 3), ("1", "0", 1),
                                                  ("1", "1", 1)])
def test_learned_cycle_matches_oracle(dev, theta40, monkeypatch, coarse32, tail, kdepth):
    """One application of the headline cycle (learned fine smoother: k0_msmooth, k0_resid or
    k0_resid32, restriction, K-cycles or V-cycles, prolongation, the fused tail) against the
    oracle's cycle with the same M and the same stored values (FP32 copies where the device's
    cycle reads them, FP64 values on the fused tail's levels): 1e-12 relative (summation
    orders differ; measured ~1e-14, while M's values taken in FP64 instead move the cycle by
    ~3e-9)."""
    prob, A, levels, lams, M = learned_hierarchy(dev, theta40, monkeypatch, coarse32, tail,
                                                 kdepth=kdepth)
    rng = np.random.default_rng(17)
    tol = 1e-12
    for _ in range(2):
        r = rng.standard_normal(dev.n)
        z = dev.mg_apply(r)
        zo = mgo.cycle(levels, 0, r, lams, M=M, kdepth=kdepth)
        err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
        print(f"coarse32 {coarse32} tail {tail}: cycle vs oracle {err:.3e}")
        assert err <= tol
    # a wrong smoother value is seen: the oracle with M's values in FP64 differs by more
    M64 = mo.m_matrix(dev.mpnn_download(2, M.nnz // 4 * mo.DEC).reshape(-1, mo.DEC),
                      mo.Graph(A.indptr, A.indices, A.data, prob.mesh.tri, dev.n // 2))
    z64 = mgo.cycle(levels, 0, r, lams, M=M64, kdepth=kdepth)
    err64 = np.linalg.norm(z - z64) / np.linalg.norm(zo)
    print(f"  with M's values in FP64: {err64:.3e}")
    assert err64 > 100 * tol


def test_learned_fcg_matches_oracle_fcg(dev, theta40, monkeypatch):
    """The flexible CG with the headline cycle, whose z'q comes from the last smoothing
    kernel's partials (k0_msmooth's epilogue), against the oracle's FCG(1) with the oracle's
    cycle: iterations within 2, solution within 1e-8."""
    prob, A, levels, lams, M = learned_hierarchy(dev, theta40, monkeypatch, "1", "0")
    b = dev.download_rhs()
    x, st = dev.solve(rtol=1e-10, max_iter=500, method=FCG)
    xo, its, hist = mgo.fcg(A, b, lambda r: mgo.cycle(levels, 0, r, lams, M=M, kdepth=3),
                            rtol=1e-10, max_iter=500)
    print(f"device {st.iterations} iterations ({st.outer_iterations} refreshes), oracle {its}")
    assert st.converged and not st.fallback and st.outer_iterations <= 2
    assert abs(st.iterations - its) <= 2, (st.iterations, its)
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo))


@pytest.mark.parametrize("smoother", [1, 0])
def test_multilevel_fcg_converges_fast(dev, smoother):
    A, xy, fixed, bptr, bcol = system(dev, hetero(120, e_seed=7, seed=9))
    _, s_cg = dev.solve(rtol=1e-10, max_iter=50000, method=cstfem_hip.METHOD_CG_FP64)
    dev.mpnn_init(0)
    dev.mpnn_set_params(learned.polynomial_start(dev.mpnn_get_params()))
    dev.mpnn_build()
    dev.mg_setup(10, 300, smoother)
    assert dev.mg_levels() >= 3
    x, s = dev.solve(rtol=1e-10, max_iter=5000, method=FCG)
    assert s.converged and not s.fallback and s.true_rel_residual < 1e-10
    assert s.iterations * 10 < s_cg.iterations, (s.iterations, s_cg.iterations)
    b = dev.download_rhs()
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10


def test_setup_order_and_arguments(dev):
    system(dev, hetero(12))
    with pytest.raises(cstfem_hip.CstfemError):
        dev.mg_setup(1, 100, 1)
    with pytest.raises(cstfem_hip.CstfemError):
        dev.mg_setup(4, 100, 7)
    dev.mg_setup(4, 100, 1)
    assert dev.mg_levels() >= 2
    dev.setup()  # a new setup drops the hierarchy
    assert dev.mg_levels() == 0
    with pytest.raises(cstfem_hip.CstfemLogicError):
        dev.mg_apply(np.ones(dev.n))


def hierarchy(dev):
    """Every level's size, aggregates and coarse matrix, and one cycle applied to a fixed r."""
    out = []
    for l in range(dev.mg_levels()):
        info = dev.mg_level_info(l)
        agg = dev.mg_download_aggregates(l) if l < dev.mg_levels() - 1 else None
        mat = dev.mg_download_matrix(l) if l >= 1 else None
        out.append((info, agg, mat))
    z = dev.mg_apply(np.random.default_rng(11).standard_normal(dev.n))
    return out, z


def assert_same_hierarchy(a, b):
    (la, za), (lb, zb) = a, b
    assert len(la) == len(lb)
    for (ia, aa, ma), (ib, ab, mb) in zip(la, lb):
        assert ia == ib
        if aa is not None:
            np.testing.assert_array_equal(aa, ab)
        if ma is not None:
            for x, y in zip(ma, mb):
                np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(za, zb)


def test_prepared_aggregation_is_the_inline_one(dev):
    """mg_prepare starts the fine level's aggregation right after the pattern (node-graph
    download on a second stream, greedy pass on a host thread) while assembly and the loads go
    on.  The set-up that takes it must build the same hierarchy, bit for bit, as one that
    aggregates inline.  A new pattern or mesh drops a started aggregation: the set-up after
    one must not use the stale result.  The big plate has more than 100,000 nodes, so the
    prepared pass runs its leftover phase on a second thread behind the root phase (bounded
    by the graph's bandwidth), against the sequential pass of the inline set-up."""
    big, small = hetero(340, e_seed=4, seed=6), hetero(90, e_seed=5, seed=7)
    assert big.mesh.num_nodes > 100_000
    ref = {}
    for name, prob in (("big", big), ("small", small)):
        pipeline.load_problem(dev, prob)
        pipeline.assemble_and_constrain(dev, prob)
        dev.setup()
        dev.mg_setup(10, 200, 1)
        ref[name] = hierarchy(dev)
    # prepared: the same hierarchy
    pipeline.load_problem(dev, big)
    pipeline.assemble_and_constrain(dev, big, multilevel=True)
    dev.setup()
    dev.mg_setup(10, 200, 1)
    assert_same_hierarchy(hierarchy(dev), ref["big"])
    # prepared twice, then the pattern is built again (assemble_and_constrain does): dropped
    dev.build_stiffness_pattern()
    dev.mg_prepare()
    dev.mg_prepare()
    pipeline.assemble_and_constrain(dev, big)
    dev.setup()
    dev.mg_setup(10, 200, 1)
    assert_same_hierarchy(hierarchy(dev), ref["big"])
    # prepared on the big plate, then another mesh: the small plate's own hierarchy
    dev.build_stiffness_pattern()
    dev.mg_prepare()
    pipeline.load_problem(dev, small)
    pipeline.assemble_and_constrain(dev, small)
    dev.setup()
    dev.mg_setup(10, 200, 1)
    assert_same_hierarchy(hierarchy(dev), ref["small"])
    # a prepared aggregation that no set-up takes is joined when the context closes
    d2 = cstfem_hip.Device(0)
    try:
        with pytest.raises(cstfem_hip.CstfemLogicError):
            d2.mg_prepare()  # no pattern yet
        pipeline.load_problem(d2, small)
        d2.build_stiffness_pattern()
        d2.mg_prepare()
    finally:
        d2.close()


def renumbered(prob, seed):
    """The same plate with its nodes in random order: the node graph's bandwidth is then
    close to the node count."""
    m = prob.mesh
    order = np.random.default_rng(seed).permutation(m.num_nodes)  # new node k = old order[k]
    new_id = np.empty_like(order)
    new_id[order] = np.arange(order.size)
    groups = {g: new_id[e].astype(np.int32) for g, e in m.groups.items()}
    mesh = dataclasses.replace(m, xy=m.xy[order].copy(), tri=new_id[m.tri].astype(np.int32),
                               groups=groups)
    return dataclasses.replace(prob, mesh=mesh)


@pytest.mark.parametrize("case", ["kirsch", "renumbered"])
def test_prepared_aggregation_on_other_graphs(dev, case):
    """The two-thread prepared pass on a graph of another shape (the Kirsch quarter plate,
    graded around the hole) and on one whose bandwidth is close to its node count (the
    leftover phase then trails the root phase by nearly the whole graph): the same hierarchy
    as the sequential inline pass."""
    prob = (pipeline.config_problem(1, 0.5) if case == "kirsch"
            else renumbered(hetero(340, e_seed=4, seed=6), 3))
    assert prob.mesh.num_nodes > 100_000
    hs = []
    for prepared in (False, True):
        pipeline.load_problem(dev, prob)
        pipeline.assemble_and_constrain(dev, prob, multilevel=prepared)
        dev.setup()
        dev.mg_setup(10, 200, 1)
        hs.append(hierarchy(dev))
    assert_same_hierarchy(hs[1], hs[0])


def with_hub(prob, nx, spread):
    """The plate with 20 extra triangles (0, a, b) fanning from node 0 to pairs of nodes
    elsewhere (overlapping elements keep the matrix SPD; node 0 ends with 22 triangles, under
    the pattern builder's 32).  Node 0 comes first in the greedy pass, so its aggregate takes
    all the fan's nodes.  spread=False: pairs of adjacent nodes along the middle row (an
    aggregate of about 30 nodes, a coarse row of about 20 blocks); spread=True: pairs two
    nodes apart on a coarse grid over the plate (an aggregate of about 70 nodes with about
    600 (member, block) pairs, a coarse row of about 80 blocks)."""
    m = prob.mesh
    W, xy = nx + 1, prob.mesh.xy
    if spread:
        s = nx // 6
        pairs = [(r * W + i, r * W + i + 2) for r in range(3, nx - 1, s)
                 for i in range(3, nx - 2, s)][:20]
    else:
        j = nx // 2
        pairs = [(j * W + i, j * W + i + 1) for i in range(nx // 2 - 10, nx // 2 + 10)]
    extra = []
    for a, b in pairs:
        d = (xy[a, 0] - xy[0, 0]) * (xy[b, 1] - xy[0, 1]) - (xy[b, 0] - xy[0, 0]) * (xy[a, 1] - xy[0, 1])
        extra.append([0, a, b] if d > 0 else [0, b, a])
    tri = np.concatenate([m.tri, np.asarray(extra, np.int32)])
    E = None if prob.E_elem is None else np.concatenate([prob.E_elem, np.ones(len(extra))])
    return dataclasses.replace(prob, mesh=dataclasses.replace(m, tri=tri), E_elem=E)


def coarse_degrees(dev, level):
    rp, cols, _ = dev.mg_download_matrix(level)
    return np.diff(rp)[0::3] // 3


def test_hub_hierarchy_matches_oracle(dev):
    """An aggregate of more (member, block) pairs than the wavefront coarse-pattern kernel's
    register slots hold (one lane then inserts them serially) and a coarse row of more than
    64 blocks (two blocks per lane in the Galerkin kernel), against the oracle: aggregates
    exact, the coarse matrix to 1e-12, one cycle to 1e-9.  Two levels: the hub's aggregate
    would be too large for a third."""
    A, xy, fixed, bptr, bcol = system(dev, with_hub(hetero(40, e_seed=8, seed=3), 40, True))
    dev.mg_setup(2, 1000, 1, 1)
    levels = mgo.setup(A, xy, fixed, max_levels=2, coarse_max=1000, ptr0=bptr, col0=bcol)
    assert dev.mg_levels() == len(levels) == 2
    agg0 = dev.mg_download_aggregates(0)
    np.testing.assert_array_equal(agg0, levels[0].agg)
    assert np.bincount(agg0, weights=np.diff(bptr)).max() > 512
    assert coarse_degrees(dev, 1).max() > 64
    nl = dev.mg_level_info(1)[0]
    rp, cols, vals = dev.mg_download_matrix(1)
    Ad = sp.csr_matrix((vals, cols, rp), shape=(nl, nl))
    assert abs(Ad - levels[1].A).max() <= 1e-12 * abs(levels[1].A).max()
    lams = [dev.mg_level_info(l)[3] for l in range(2)]
    r = np.random.default_rng(5).standard_normal(dev.n)
    zo = mgo.cycle(levels, 0, r, lams, kdepth=1)
    assert np.linalg.norm(dev.mg_apply(r) - zo) <= 1e-9 * np.linalg.norm(zo)


def test_setup_lane_forms_match_wavefront_forms(dev, monkeypatch):
    """On a level of many small aggregates the Gram-Schmidt and coarse-pattern kernels run one
    lane per aggregate and hand the aggregates above a lane's registers (more than 10 nodes,
    more than 16 coarse blocks) to the wavefront forms through a list.  The hierarchy must be
    the one the wavefront forms alone build (CSTFEM_MG_SETUP_LANES=0), bit for bit."""
    prob = with_hub(hetero(1100, e_seed=8, seed=3), 1100, False)
    out = []
    for lanes in ("1", "0"):
        monkeypatch.setenv("CSTFEM_MG_SETUP_LANES", lanes)
        pipeline.load_problem(dev, prob)
        pipeline.assemble_and_constrain(dev, prob)
        dev.setup()
        dev.mg_setup(10, 200, 1)
        out.append(hierarchy(dev))
    assert dev.mg_level_info(0)[1] >= 65536  # the lane forms ran on level 0
    assert np.bincount(dev.mg_download_aggregates(0)).max() > 10
    assert coarse_degrees(dev, 1).max() > 16
    assert_same_hierarchy(out[0], out[1])


def test_coarse_level_above_one_k_cycle_grid(dev):
    """A first coarse level larger than the K-cycle's reduction grid covers (2048 workgroups x
    256 threads = 524288 DOFs): every K-cycle kernel must stride over the whole level."""
    system(dev, hetero(1300, e_seed=7, seed=9))
    dev.mg_setup(10, 300, 1)
    assert dev.mg_level_info(1)[0] > 524288
    x, s = dev.solve(rtol=1e-10, max_iter=2000, method=FCG)
    assert s.converged and not s.fallback and s.true_rel_residual < 1e-10
    assert s.iterations < 150, s.iterations


@pytest.mark.parametrize("kdepth", [1, 2])
def test_fused_tail_matches_separate_kernels(dev, monkeypatch, kdepth):
    """The smallest V-cycle levels run in one single-workgroup kernel; switching that off
    (CSTFEM_MG_TAIL=0, read at set-up) must give the same cycle up to summation order."""
    system(dev, hetero(60))
    dev.mg_setup(8, 20, 1, kdepth)
    sizes = [dev.mg_level_info(l)[0] for l in range(dev.mg_levels())]
    assert len(sizes) >= kdepth + 3 and sizes[kdepth + 1] <= 3072, sizes
    r = np.random.default_rng(3).standard_normal(dev.n)
    z_tail = dev.mg_apply(r)
    monkeypatch.setenv("CSTFEM_MG_TAIL", "0")
    dev.mg_setup(8, 20, 1, kdepth)
    z_kern = dev.mg_apply(r)
    assert np.linalg.norm(z_tail - z_kern) <= 1e-12 * np.linalg.norm(z_kern)
    assert np.linalg.norm(z_kern) > 0


@pytest.mark.parametrize("smoother", [1, 0])
def test_packed_tail_is_bit_identical(dev, theta40, monkeypatch, smoother):
    """A one-level tail over the dense coarsest (here 867 DOFs over 108, the shape of the
    4M-DOF plate's 972 over 108) runs k_tail1 on operands packed at set-up; with the pack
    switched off (CSTFEM_MG_TAIL_PACK=0) k_tail reads them level by level.  Same arithmetic in
    the same order: one cycle and a whole flexible CG solve are bit-identical, and both match
    the separate kernels (CSTFEM_MG_TAIL=0) up to summation order (FP64 values) or to the
    rounding of the FP32 copies those kernels read with the learned smoother."""
    system(dev, hetero(150))
    if smoother == 0:
        dev.mpnn_set_params(theta40)
        dev.mpnn_build()
    r = np.random.default_rng(12).standard_normal(dev.n)
    out = {}
    for pack, tail in (("1", "1"), ("0", "1"), ("1", "0")):
        monkeypatch.setenv("CSTFEM_MG_TAIL_PACK", pack)
        monkeypatch.setenv("CSTFEM_MG_TAIL", tail)
        dev.mg_setup(8, 120, smoother, 1)
        sizes = [dev.mg_level_info(l)[0] for l in range(dev.mg_levels())]
        assert sizes[-2:] == [867, 108], sizes
        first, packed = dev.mg_tail()
        assert (first, packed) == ((2, 1) if pack + tail == "11" else (2, 0) if tail == "1" else (-1, 0))
        z = dev.mg_apply(r)
        dev.mg_apply(r)  # the second call reuses the pack
        x, s = dev.solve(rtol=1e-10, max_iter=500, method=FCG)
        assert s.converged and s.true_rel_residual < 1e-10 and not s.fallback
        out[pack + tail] = (z, x, s.iterations)
    np.testing.assert_array_equal(out["11"][0], out["01"][0])
    np.testing.assert_array_equal(out["11"][1], out["01"][1])
    assert out["11"][2] == out["01"][2]
    # the separate kernels read the FP32 copies of the tail levels' values where the learned
    # smoother's cycle keeps them (the fused tails read FP64): relative rounding 2^-24 per entry
    z_sep = out["10"][0]
    tol = 1e-12 if smoother == 1 else 1e-6
    assert np.linalg.norm(out["11"][0] - z_sep) <= tol * np.linalg.norm(z_sep)
    assert abs(out["11"][2] - out["10"][2]) <= 1


@pytest.mark.parametrize("smoother,kdepth", [(0, 3), (0, 1), (1, 3)])
def test_fused_cycle_kernels_are_bit_identical(dev, theta40, monkeypatch, smoother, kdepth):
    """The fused cycle kernels (K-cycle updates folded into their readers, CSTFEM_MG_FOLD;
    residual + restriction in one kernel and the prolongation inside the first post-smoothing
    step on levels below CSTFEM_MG_FUSE_N DOFs) perform the separate kernels' arithmetic in the same
    order: one cycle, and a whole flexible CG solve, are bit-identical to CSTFEM_MG_FUSE=0,
    with FP32 coarse values (learned smoother) and FP64 ones (Chebyshev), with and without
    the fused tail below the K-cycle levels."""
    system(dev, hetero(300, e_seed=3, seed=4))
    if smoother == 0:
        dev.mpnn_set_params(theta40)
        dev.mpnn_build()
    r = np.random.default_rng(11).standard_normal(dev.n)
    out = {}
    for fuse, fuse_n, fold in (("0", "200000", "200000"), ("1", "0", "1000000000"),
                               ("1", "200000", "200000"), ("1", "100000000", "1000000000"),
                               ("1", "100000000", "0")):
        monkeypatch.setenv("CSTFEM_MG_FUSE", fuse)
        monkeypatch.setenv("CSTFEM_MG_FUSE_N", fuse_n)
        monkeypatch.setenv("CSTFEM_MG_FOLD", fold)
        dev.mg_setup(12, 500, smoother, kdepth)
        if smoother == 0:
            assert 0.5 < dev.mg_level_info(0)[3] <= 1.5  # the set-up kept the learned smoother
        sizes = [dev.mg_level_info(l)[0] for l in range(dev.mg_levels())]
        z = dev.mg_apply(r)
        x, st = dev.solve(rtol=1e-10, max_iter=2000, method=FCG)
        assert st.converged and not st.fallback and st.true_rel_residual < 1e-10
        out[(fuse, fuse_n, fold)] = (z, x, st.iterations)
        print(f"fuse {fuse} below {fuse_n} DOFs, fold {fold}: levels {sizes}, "
              f"{st.iterations} iterations")
    z0, x0, it0 = out[("0", "200000", "200000")]
    assert np.linalg.norm(z0) > 0
    for key, (z, x, it) in out.items():
        np.testing.assert_array_equal(z, z0, err_msg=str(key))
        np.testing.assert_array_equal(x, x0, err_msg=str(key))
        assert it == it0, key


def test_trained_smoother_beats_untrained_and_chebyshev(dev):
    theta, losses = learned.train(dev, learned.training_plate(64), steps=300)
    assert losses[-1] < 0.8 * losses[0]
    system(dev, hetero(200, e_seed=7, seed=9))
    its = {}
    for name, th, smoother in (("trained", theta, 0), ("untrained", learned.polynomial_start(theta), 0),
                               ("chebyshev", None, 1)):
        dev.setup()
        if th is not None:
            dev.mpnn_set_params(th)
            dev.mpnn_build()
        dev.mg_setup(12, 500, smoother, 3)
        if smoother == 0:
            assert 0.5 < dev.mg_level_info(0)[3] <= 1.5, name  # accepted by the set-up check
        _, s = dev.solve(rtol=1e-10, max_iter=2000, method=FCG)
        assert s.converged and not s.fallback and s.true_rel_residual < 1e-10, name
        its[name] = s.iterations
    assert its["trained"] < its["untrained"] and its["trained"] < its["chebyshev"], its


@pytest.mark.parametrize("bad", ["nan", "indefinite"])
def test_bad_learned_smoother_falls_back_to_chebyshev_cycle(dev, bad):
    A, xy, fixed, bptr, bcol = system(dev, hetero(120, e_seed=7, seed=9))
    dev.mpnn_init(0)
    th = dev.mpnn_get_params()
    if bad == "nan":
        th[learned.BD] = np.nan      # M = NaN: breakdown at the first iteration
    else:
        th[learned.BD + 2] = -1.0    # M = diag(+D^-1, -D^-1): indefinite
    dev.mpnn_set_params(th)
    dev.mpnn_build()
    dev.mg_setup(10, 300, 0)
    # the set-up rejects a smoother whose lambda_max(M A) is above 1.5 (or not finite) and
    # uses Chebyshev from the start; one that passes and then fails hands over at run time
    rejected = not (dev.mg_level_info(0)[3] <= 1.5)
    x, s = dev.solve(rtol=1e-10, max_iter=5000, method=FCG)
    assert s.converged and s.true_rel_residual < 1e-10, (s.fallback, s.iterations)
    assert (rejected and s.fallback == 0) or s.fallback == 2, (rejected, s.fallback)
    assert s.iterations < 600, s.iterations
    b = dev.download_rhs()
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10


def test_fp32_coarse_values_match_fp64_cycle(dev, monkeypatch):
    """The learned smoother's cycle reads FP32 copies of the coarse matrices by default;
    CSTFEM_MG_COARSE32=0 (read at set-up) keeps FP64.  One application differs by the FP32
    rounding of the coarse values only."""
    system(dev, hetero(60))
    dev.mpnn_init(0)
    dev.mpnn_set_params(learned.polynomial_start(dev.mpnn_get_params()))
    dev.mpnn_build()
    r = np.random.default_rng(4).standard_normal(dev.n)
    dev.mg_setup(8, 20, 0, 2)
    z32 = dev.mg_apply(r)
    monkeypatch.setenv("CSTFEM_MG_COARSE32", "0")
    dev.mg_setup(8, 20, 0, 2)
    z64 = dev.mg_apply(r)
    err = np.linalg.norm(z32 - z64) / np.linalg.norm(z64)
    assert 0 < err < 1e-5, err

"""CPU tests of the oracle (the checker), pinned against things it does not
compute itself: the reference's own Plummer sampler and published tree
counts, closed forms, brute force, and structural invariants."""
import itertools

import numpy as np
import pytest


# ---- reference sampler and published known answer ----------------------------

def test_plummer_restatement_matches_reference_sampler(oracle):
    """or_plummer_ref is bit-exact against code/plummer.h compiled into oracle/_ref."""
    ref = oracle.ref_host()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for n in (64, 1000, 4096):
        m1, p1, v1 = oracle.plummer_ref(n)
        m2, p2, v2 = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
        assert ref.ref_plummer(n, 19810614, oracle._ptr(m2), oracle._ptr(p2), oracle._ptr(v2)) == 0
        assert np.array_equal(m1, m2)
        assert np.array_equal(p1.ravel(), p2)
        assert np.array_equal(v1.ravel(), v2)


def test_tree_counts_match_published_known_answer(oracle):
    """treebuild/readme:32-40: Plummer N=2e7, seed 19810614, NLEAF=16 ->
    nb_leaf 20000000, nnodes 596463, nleaves 4041848, ncells 4638311, nlevels 15."""
    pos, _ = oracle.plummer_ref_f4(20_000_000, 19810614, 100.0)
    t = oracle.build_tree(pos, 16)
    assert t.nb_leaf == 20_000_000
    assert t.nnodes == 596463
    assert t.nleaves == 4041848
    assert t.ncells == 4638311
    assert t.nlevels == 15
    assert t.capped == 0


# ---- root box ------------------------------------------------------------------

def test_bounding_box_hand_case(oracle):
    # extent (0,0,0)..(3,1,1): centre (1.5,.5,.5), half extent 1.5 -> hsize 2,
    # centre already a multiple of 2/2^20
    pos = np.array([[0, 0, 0, 1], [3, 1, 1, 1]], np.float32)
    box, mn, mx = oracle.bounding_box(pos)
    assert list(box) == [1.5, 0.5, 0.5, 2.0]
    assert list(mn) == [0, 0, 0] and list(mx) == [3, 1, 1]


def test_bounding_box_properties(oracle):
    rng = np.random.default_rng(3)
    for scale, shift in ((1.0, 0.0), (1e-3, 5.0), (40.0, -7.0)):
        pos = np.zeros((1000, 4), np.float32)
        pos[:, :3] = rng.normal(shift, scale, (1000, 3))
        box, mn, mx = oracle.bounding_box(pos)
        c, h = box[:3].astype(np.float64), float(box[3])
        assert np.array_equal(mn, pos[:, :3].min(0)) and np.array_equal(mx, pos[:, :3].max(0))
        half = float(((mx.astype(np.float64) - mn) / 2).max())
        assert h >= half > h / 2                       # smallest power of two >= extent
        assert np.log2(h) == np.round(np.log2(h))
        q = h / 2 ** 20                                 # centre quantised to hsize/2^20
        assert np.all(np.abs(c / q - np.round(c / q)) == 0)
        # within one quantum of the float32 midpoint (code/buildTree.cu:145 is float)
        mid = (mx.astype(np.float64) + mn) / 2
        assert np.all(np.abs(c - mid) < q * 1.0001 + np.spacing(np.abs(mid).astype(np.float32)))


# ---- octree ----------------------------------------------------------------------

def test_recursive_partition_equals_geometric_key_sort(oracle, bt):
    """Two routes to the tree order: recursive stable partition by ChildBox
    octants vs a stable sort of each particle's own octant-digit key."""
    pos, _ = bt.plummer(20000, seed=5)
    box, _, _ = oracle.bounding_box(pos)
    t = oracle.build_tree(pos, 16, box)
    keys = oracle.geo_keys(pos, box)
    assert np.array_equal(t.order, np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("nleaf", [16, 24, 32, 48, 64])
def test_tree_invariants(oracle, bt, nleaf):
    pos, _ = bt.clustered(20000, seed=9)
    box, _, _ = oracle.bounding_box(pos)
    t = oracle.build_tree(pos, nleaf, box)
    keys = oracle.geo_keys(pos, box)[t.order]
    cells = t.cells.astype(np.int64)
    level = cells[:, 0] >> 27
    leaf = cells[:, 1] == 0xFFFFFFFF
    b, e = cells[:, 2], cells[:, 3]
    np_ = e - b
    assert np_[leaf].sum() == len(pos)                  # SWtest/treebuild.cu:1477
    assert (np_[leaf] <= nleaf).all() and (np_[~leaf] > nleaf).all()
    assert (np.diff(level) >= 0).all()                  # level order
    for L in range(1, t.nlevels + 1):
        lo, hi = t.level_range[L]
        assert (level[lo:hi + 1] == L).all()
    first = cells[:, 1] & 0x1FFFFFFF
    nch = (cells[:, 1] >> 29) + 1
    for c in np.where(~leaf)[0]:
        kids = np.arange(first[c], first[c] + nch[c])
        assert (level[kids] == level[c] + 1).all()
        # children partition the parent's range
        rng_ = sorted(zip(b[kids], e[kids]))
        assert rng_[0][0] == b[c] and rng_[-1][1] == e[c]
        assert all(rng_[i][1] == rng_[i + 1][0] for i in range(len(rng_) - 1))
        # nodes first, then leaves (code/buildTree.cu:393-400)
        lk = leaf[kids]
        assert not (lk[:-1] & ~lk[1:]).any()
    # every particle of a cell shares the cell's octant-digit prefix
    for c in range(0, len(cells), 97):
        sh = np.uint64(63 - 3 * int(level[c]))
        pre = keys[b[c]:e[c]] >> sh
        assert (pre == pre[0]).all()


def test_depth_cap_on_coincident_particles(oracle):
    pos = np.zeros((40, 4), np.float32)
    pos[:, 3] = 1.0 / 40
    pos[:20, :3] = 0.25  # 20 coincident particles > NLEAF: no split can separate them
    pos[20:, 0] = np.linspace(-1, 1, 20)
    t = oracle.build_tree(pos, 16)
    assert t.capped == 1 and t.nb_leaf == 40 and t.nlevels == 21


# ---- Peano-Hilbert keys -----------------------------------------------------------

@pytest.mark.parametrize("nbits", [1, 2, 3])
def test_hilbert_key_bijective_and_face_adjacent(oracle, nbits):
    """code/makeGroups.cu:19-106 is a Hilbert curve: a bijection on the
    (2^nbits)^3 cells whose consecutive keys are face neighbours."""
    keys = {}
    for x, y, z in itertools.product(range(2 ** nbits), repeat=3):
        keys[oracle.hilbert_key(x, y, z, nbits)] = (x, y, z)
    assert sorted(keys) == list(range(8 ** nbits))
    walk = [keys[k] for k in range(8 ** nbits)]
    assert all(sum(abs(a - b) for a, b in zip(walk[i], walk[i + 1])) == 1
               for i in range(len(walk) - 1))


def test_hilbert_key_golden(oracle):
    # top digit only (lower bits zero): output digit table of makeGroups.cu:28
    top = 1 << 20
    table = [0, 1, 7, 6, 3, 2, 4, 5]
    for idx in range(8):
        x, y, z = (idx >> 2) & 1, (idx >> 1) & 1, idx & 1
        k = oracle.hilbert_key(x * top, y * top, z * top, 21)
        assert k >> 60 == table[idx]
    assert oracle.hilbert_key(0, 0, 0, 21) == 0


# ---- multipoles ---------------------------------------------------------------------

def _cell(pos):
    return np.array([[1 << 27, 0xFFFFFFFF, 0, len(pos)]], np.uint32)


def test_multipole_closed_forms(oracle):
    # one particle: com = x, Q = 0, l = 1e-6 floor, s = 0
    p = np.array([[0.3, -0.2, 0.1, 0.5]], np.float32)
    size, mono, q0, q1, ls = oracle.multipoles(_cell(p), p, 0.5)
    assert np.array_equal(mono[0], p[0]) and not q0.any() and not q1.any()
    assert ls[0, 0] == np.float32(1e-6) and ls[0, 1] == 0
    # two equal masses at x = +-a: com 0, Q_xx = a^2, l = 2a, cellOp = 2a/theta
    a = 0.25
    p = np.array([[a, 0, 0, 1], [-a, 0, 0, 1]], np.float32)
    size, mono, q0, q1, ls = oracle.multipoles(_cell(p), p, 0.5)
    assert mono[0, 3] == 2 and np.allclose(mono[0, :3], 0)
    assert q0[0, 0] == np.float32(a * a) and not q0[0, 1:].any() and not q1.any()
    assert ls[0, 0] == 2 * a and size[0, 3] == np.float32((2 * a / 0.5) ** 2)


def test_multipole_parallel_axis_and_psd(oracle, bt):
    """Q of a node equals the mass-weighted children Q plus the spread of the
    children's centres (parallel-axis theorem); Q is positive semi-definite."""
    pos, _ = bt.plummer(5000, seed=4)
    t = oracle.build_tree(pos, 16)
    pt = pos[t.order]
    size, mono, q0, q1, ls = oracle.multipoles(t.cells, pt, 0.5)
    Q = np.zeros((len(t.cells), 3, 3))
    Q[:, 0, 0], Q[:, 1, 1], Q[:, 2, 2] = q0[:, 0], q0[:, 1], q0[:, 2]
    Q[:, 0, 1] = Q[:, 1, 0] = q0[:, 3]
    Q[:, 0, 2] = Q[:, 2, 0] = q1[:, 0]
    Q[:, 1, 2] = Q[:, 2, 1] = q1[:, 1]
    assert (np.linalg.eigvalsh(Q).min(axis=1) >= -1e-7 * ls[:, 0] ** 2).all()
    cells = t.cells.astype(np.int64)
    for c in np.where(cells[:, 1] != 0xFFFFFFFF)[0][:200]:
        f, n = cells[c, 1] & 0x1FFFFFFF, (cells[c, 1] >> 29) + 1
        kids = np.arange(f, f + n)
        m = mono[kids, 3].astype(np.float64)
        x = mono[kids, :3].astype(np.float64)
        M = m.sum()
        com = (m[:, None] * x).sum(0) / M
        d = x - com
        Qp = ((m[:, None, None] * (Q[kids] + d[:, :, None] * d[:, None, :])).sum(0)) / M
        assert np.allclose(com, mono[c, :3], atol=1e-6 * ls[c, 0] + 1e-7)
        assert np.allclose(Qp, Q[c], atol=2e-6 * ls[c, 0] ** 2)


def test_reference_fp32_moment_arithmetic_loses_digits(oracle):
    """The reference's FP32 products (code/computeMultipoles.cu:113-158) lose
    most digits of Q for a small cell far from the origin; the FP64 mode does
    not.  This is why the oracle's default sums moments in FP64."""
    rng = np.random.default_rng(1)
    p = np.zeros((16, 4), np.float32)
    p[:, :3] = 10.0 + 1e-3 * rng.random((16, 3))
    p[:, 3] = 1.0 / 16
    exact = np.cov(p[:, :3].astype(np.float64).T, bias=True)[0, 0]
    q64 = oracle.multipoles(_cell(p), p, 0.5, mode=0)[2][0, 0]
    q32 = oracle.multipoles(_cell(p), p, 0.5, mode=1)[2][0, 0]
    assert abs(q64 - exact) < 1e-6 * exact
    assert abs(q32 - exact) > 1e-2 * exact


# ---- groups ---------------------------------------------------------------------------

def test_groups_closed_form(oracle):
    # masked level-1 digits: runs of length 5, 3, 130 -> chop by 64
    keys = np.array([0] * 5 + [1 << 60] * 3 + [3 << 60] * 130, np.uint64)
    g = oracle.make_groups(keys, 1, 64)
    assert g.tolist() == [[0, 5], [5, 3], [8, 64], [72, 64], [136, 2]]
    # lower digits do not split at level_split 1 but do at 2
    keys = np.array([0, 1 << 57, 2 << 57], np.uint64)
    assert oracle.make_groups(keys, 1, 64).tolist() == [[0, 3]]
    assert oracle.make_groups(keys, 2, 64).tolist() == [[0, 1], [1, 1], [2, 1]]


# ---- walk ------------------------------------------------------------------------------

def _pipeline(oracle, pos, theta, eps, nleaf=16, ncrit=64, acc_type=0):
    return oracle.reference_pipeline(pos, nleaf, ncrit, theta, eps, 5, acc_type)


def test_walk_theta_to_zero_is_direct_sum(oracle, bt):
    pos, _ = bt.plummer(2000, seed=2)
    r = _pipeline(oracle, pos, 1e-4, 0.01, acc_type=1)
    assert r["stats"][2] == 0  # no approximations left
    d = oracle.direct_fp64(r["pos_ph"], np.arange(len(pos), dtype=np.int32), 0.01)
    rel = np.linalg.norm(r["acc"][:, :3] - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
    assert rel.max() < 1e-5
    assert np.allclose(r["acc"][:, 3], d[:, 3], rtol=1e-5)


@pytest.mark.parametrize("nleaf", [16, 64])
def test_leaf_groups_tile_the_particles(oracle, bt, nleaf):
    """Leaf groups (code/buildTree.cu:954-958): one group per leaf, in tree
    order, tiling [0, n) with at most NLEAF particles each; a walk over them
    at theta -> 0 is the direct sum, at theta 0.5 it stays within the
    quadrupole error of the PH-group walk."""
    pos, _ = bt.plummer(3000, seed=4)
    box, _, _ = oracle.bounding_box(pos)
    t = oracle.build_tree(pos, nleaf, box)
    g = oracle.leaf_groups(t.cells)
    nleaves = int((t.cells[:, 1] == 0xFFFFFFFF).sum())
    assert len(g) == nleaves
    assert g[0, 0] == 0 and g[-1, 0] + g[-1, 1] == len(pos)
    assert (g[:-1, 0] + g[:-1, 1] == g[1:, 0]).all()
    assert (g[:, 1] >= 1).all() and (g[:, 1] <= nleaf).all()
    pos_tree = pos[t.order]
    d = oracle.direct_fp64(pos_tree, np.arange(len(pos), dtype=np.int32), 0.01)
    for theta, tol in ((1e-4, 1e-5), (0.5, 5e-3)):
        size, mono, q0, q1, _ = oracle.multipoles(t.cells, pos_tree, theta)
        acc, stats, _ = oracle.walk(g, t.level_range, pos_tree, pos_tree, t.cells, size, mono, q0,
                                    q1, 0.01)
        rel = np.linalg.norm(acc[:, :3] - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
        assert np.percentile(rel, 99) < tol, (theta, np.percentile(rel, 99))
        if theta < 1e-3:
            assert stats[2] == 0


@pytest.mark.parametrize("nleaf,ncrit", [(16, 64), (32, 32), (64, 48)])
def test_tree_groups_are_the_topmost_small_cells(oracle, bt, nleaf, ncrit):
    """Tree groups: from every leaf, climb while the parent holds <= ncrit
    particles; the cells reached, ordered by first particle, are the groups
    (an independent derivation of oracle.tree_groups).  They tile [0, N)."""
    pos, _ = bt.clustered(5000, seed=6)
    box, _, _ = oracle.bounding_box(pos)
    t = oracle.build_tree(pos, nleaf, box)
    cells = t.cells.astype(np.int64)
    np_ = cells[:, 3] - cells[:, 2]
    level = cells[:, 0] >> 27
    parent = cells[:, 0] & 0x07FFFFFF
    top = set()
    for c in np.where(cells[:, 1] == 0xFFFFFFFF)[0]:
        while level[c] > 1 and np_[parent[c]] <= ncrit:
            c = parent[c]
        top.add(int(c))
    want = sorted((int(cells[c, 2]), int(np_[c])) for c in top)
    g = oracle.tree_groups(t.cells, ncrit)
    assert [tuple(x) for x in g.tolist()] == want
    assert g[0, 0] == 0 and g[-1, 0] + g[-1, 1] == len(pos)
    assert (g[:-1, 0] + g[:-1, 1] == g[1:, 0]).all()
    assert (g[:, 1] <= max(ncrit, nleaf)).all()


def test_walk_two_body_closed_form(oracle):
    m1, m2, r, eps = 0.7, 0.3, 0.4, 0.05
    pos = np.array([[0, 0, 0, m1], [r, 0, 0, m2]], np.float32)
    res = _pipeline(oracle, pos, 0.5, eps, acc_type=1)
    s = (r * r + eps * eps) ** -1.5
    acc = {round(float(p[0]), 3): a for p, a in zip(res["pos_ph"], res["acc"])}
    assert np.isclose(acc[0.0][0], m2 * r * s, rtol=1e-6)
    assert np.isclose(acc[0.4][0], -m1 * r * s, rtol=1e-6)
    # potential includes the self term -m_i/eps (code/computeForces.cu:64-72)
    assert np.isclose(acc[0.0][3], -m1 / eps - m2 / np.sqrt(r * r + eps * eps), rtol=1e-6)


def test_quadrupole_error_scales_as_third_power(oracle):
    """P2C of a 4-particle cell vs the exact sum: the truncation error
    of a monopole+quadrupole expansion falls as (a/d)^3 (next term: octupole)."""
    a = 0.05
    errs = []
    for d in (1.0, 2.0, 4.0):
        # 4 particles: cells with < 3 particles are always opened (:284-285)
        pts = np.array([[a, a / 3, 0, 0.3], [-a, -a / 3, 0, 0.3], [a / 2, -a, a / 4, 0.2],
                        [-a / 2, a, -a / 4, 0.2]], np.float64)
        cell = _cell(pts.astype(np.float32))
        size, mono, q0, q1, ls = oracle.multipoles(cell, pts.astype(np.float32), 0.5)
        target = np.array([[d, 0.3 * d, 0.1 * d, 1e-9]], np.float32)
        lr = np.zeros(64, np.int32)
        lr[2:4] = [0, 0]
        acc, _, _ = oracle.walk(np.array([[0, 1]], np.int32), lr, target, pts.astype(np.float32),
                                cell, size * 0, mono, q0, q1, 0.0, acc_type=1)
        x = target[0, :3].astype(np.float64)
        ex = sum(p[3] * (p[:3] - x) / np.linalg.norm(p[:3] - x) ** 3 for p in pts)
        errs.append(np.linalg.norm(acc[0, :3] - ex) / np.linalg.norm(ex))
    assert errs[0] / errs[1] > 6 and errs[1] / errs[2] > 6  # ~8 per doubling


def test_monopole_mode_is_a_point_mass_with_second_order_error(oracle):
    """Monopole-only P2C (the reference without QUADRUPOLE,
    code/computeForces.cu:198-225) is the cell's mass at its centre of mass:
    equal to that closed form, and its truncation error falls as (a/d)^2."""
    a = 0.05
    errs = []
    for d in (1.0, 2.0, 4.0):
        pts = np.array([[a, a / 3, 0, 0.3], [-a, -a / 3, 0, 0.3], [a / 2, -a, a / 4, 0.2],
                        [-a / 2, a, -a / 4, 0.2]], np.float64)
        cell = _cell(pts.astype(np.float32))
        size, mono, q0, q1, ls = oracle.multipoles(cell, pts.astype(np.float32), 0.5)
        target = np.array([[d, 0.3 * d, 0.1 * d, 1e-9]], np.float32)
        lr = np.zeros(64, np.int32)
        lr[2:4] = [0, 0]
        acc, st, _ = oracle.walk(np.array([[0, 1]], np.int32), lr, target,
                                 pts.astype(np.float32), cell, size * 0, mono, q0, q1, 0.0,
                                 acc_type=1, monopole=True)
        assert st[2] == 1  # one P2C
        x = target[0, :3].astype(np.float64)
        com, M = mono[0, :3].astype(np.float64), float(mono[0, 3])
        point = M * (com - x) / np.linalg.norm(com - x) ** 3
        assert np.allclose(acc[0, :3], point, rtol=1e-6)
        ex = sum(p[3] * (p[:3] - x) / np.linalg.norm(p[:3] - x) ** 3 for p in pts)
        errs.append(np.linalg.norm(acc[0, :3] - ex) / np.linalg.norm(ex))
    assert 3 < errs[0] / errs[1] < 6 and 3 < errs[1] / errs[2] < 6  # ~4 per doubling


def test_direct_fp64_against_long_double(oracle, bt):
    pos, _ = bt.plummer(300, seed=8)
    idx = np.arange(300, dtype=np.int32)
    a = oracle.direct_fp64(pos, idx, 0.01)
    b = oracle.direct_fp64(pos, idx, 0.01, longdouble=True)
    assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_energies_two_body(oracle):
    m1, m2, r, eps = 0.6, 0.4, 0.5, 0.05
    pos = np.array([[0, 0, 0, m1], [r, 0, 0, m2]], np.float32)
    vel = np.array([[0, 0.3, 0, m1], [0, -0.45, 0, m2]], np.float32)
    acc = oracle.direct_fp64(pos, np.arange(2, dtype=np.int32), eps).astype(np.float32)
    ek, ep = oracle.energies(pos, vel, acc, eps)
    assert np.isclose(ek, 0.5 * m1 * 0.3 ** 2 + 0.5 * m2 * 0.45 ** 2, rtol=1e-6)
    assert np.isclose(ep, -m1 * m2 / np.sqrt(r * r + eps * eps), rtol=1e-5)


def test_walk_stats_are_list_lengths(oracle, bt):
    pos, _ = bt.plummer(3000, seed=6)
    r = _pipeline(oracle, pos, 0.6, 0.01)
    acc, st, pot, cnt = oracle.walk(r["groups"], r["tree"].level_range, r["pos_ph"],
                                    r["pos_tree"], r["tree"].cells, r["size"], r["mono"],
                                    r["q0"], r["q1"], 0.01, counts=True)
    npg = r["groups"][:, 1].astype(np.int64)
    assert st[0] == (cnt[:, 1] * npg).sum() and st[2] == (cnt[:, 0] * npg).sum()
    assert st[1] == cnt[:, 1].max() and st[3] == cnt[:, 0].max()
    assert np.isclose(pot, 0.5 * (r["pos_ph"][:, 3].astype(np.float64) * acc[:, 3]).sum())


# ---- the opening rule, pinned independently of the oracle's own arithmetic ----

def _cell_geometry_fp64(cells, pos_tree, theta):
    """Per cell, from its particles in FP64: centre of mass, particle bbox,
    opening radius op = l/theta + s with l = max(2 * max half extent, 1e-6)
    and s = |bbox centre - com| (code/computeMultipoles.cu:168-177)."""
    p = pos_tree[:, :3].astype(np.float64)
    m = pos_tree[:, 3].astype(np.float64)
    cm = np.concatenate([[0.0], np.cumsum(m)])
    cx = np.concatenate([np.zeros((1, 3)), np.cumsum(p * m[:, None], axis=0)])
    beg = cells[:, 2].astype(np.int64)
    end = cells[:, 3].astype(np.int64)
    M = cm[end] - cm[beg]
    com = (cx[end] - cx[beg]) / M[:, None]
    bmin = np.array([p[b:e].min(axis=0) for b, e in zip(beg, end)])
    bmax = np.array([p[b:e].max(axis=0) for b, e in zip(beg, end)])
    l = np.maximum(2.0 * (0.5 * (bmax - bmin)).max(axis=1), 1e-6)
    s = np.linalg.norm(0.5 * (bmax + bmin) - com, axis=1)
    return com, l / theta + s, l / theta, end - beg


def _check_lists(alist, leaves, gmin, gmax, cells, com, op, npc, n, tol=1e-5):
    """The rule as the walk must apply it, in FP64: ds = |com - clamp(com,
    group box)| (code/computeForces.cu:30-54); a cell is accepted (P2C) iff
    ds >= op and it holds >= 3 particles (:284-285); accepted cells are
    maximal (their parent was opened); accepted cells and opened leaves tile
    the particles exactly once.  Returns the list of violations."""
    bad = []
    def ds(c):
        return np.linalg.norm(com[c] - np.clip(com[c], gmin, gmax), axis=-1)
    a = np.asarray(alist, np.int64)
    lf = np.asarray(leaves, np.int64)
    if len(a):
        if (npc[a] < 3).any():
            bad.append("accepted cell with < 3 particles")
        if (ds(a) < op[a] * (1 - tol)).any():
            bad.append("accepted cell inside its opening radius")
        level = cells[a, 0] >> 27
        par = (cells[a, 0] & 0x07FFFFFF).astype(np.int64)[level > 1]
        if len(par) and not ((ds(par) < op[par] * (1 + tol)) | (npc[par] < 3)).all():
            bad.append("accepted cell whose parent was acceptable")
    if len(lf) and not ((ds(lf) < op[lf] * (1 + tol)) | (npc[lf] < 3)).all():
        bad.append("opened leaf outside its opening radius")
    rng = np.concatenate([cells[a, 2:4], cells[lf, 2:4]]).astype(np.int64)
    rng = rng[np.argsort(rng[:, 0])]
    if not (rng[0, 0] == 0 and rng[-1, 1] == n and (rng[1:, 0] == rng[:-1, 1]).all()):
        bad.append("lists do not tile the particles")
    return bad


def test_opening_rule_fp64_clamp_distance(oracle, bt):
    """The oracle's interaction lists (or_walk_lists, the traversal of
    code/computeForces.cu:229-463) obey an FP64 statement of the opening rule
    computed from the particles alone.  Each planted mistake -- no < 3
    particles rule, the group as a point, no s term -- breaks the check, and
    the sample holds cells on both sides of l/theta + s that tell them apart."""
    n, theta = 20000, 0.5
    pos, _ = bt.plummer(n, seed=41)
    r = _pipeline(oracle, pos, theta, 0.01)
    cells = r["tree"].cells
    com, op, op_nos, npc = _cell_geometry_fp64(cells, r["pos_tree"], theta)
    rng = np.random.default_rng(3)
    sel = rng.choice(len(r["groups"]), size=24, replace=False)
    found = {1: False, 2: False, 3: False}
    between = 0
    for g in sel:
        pb, npg = r["groups"][g]
        gp = r["pos_ph"][pb:pb + npg, :3].astype(np.float64)
        gmin, gmax = gp.min(axis=0), gp.max(axis=0)
        a, lf = oracle.walk_lists(g, r["groups"], r["tree"].level_range, r["pos_ph"], cells,
                                  r["size"])
        assert _check_lists(a, lf, gmin, gmax, cells, com, op, npc, n) == [], g
        d = np.linalg.norm(com - np.clip(com, gmin, gmax), axis=1)
        between += int(((d >= op_nos) & (d < op)).sum())
        for mut in (1, 2, 3):
            am, lm = oracle.walk_lists(g, r["groups"], r["tree"].level_range, r["pos_ph"], cells,
                                       r["size"], ls=r["ls"], theta=theta, mutate=mut)
            if _check_lists(am, lm, gmin, gmax, cells, com, op, npc, n):
                found[mut] = True
    assert between > 0
    assert all(found.values()), found


# ---- the Treecode<double> oracle -------------------------------------------

def test_fp64_oracle_dyadic_inputs_build_the_float_tree(oracle, bt):
    """On inputs whose coordinates are short dyadic rationals, every ChildBox
    and quantisation step is exact in both precisions: the FP64 box, tree,
    order and keys equal the FP32 ones."""
    rng = np.random.default_rng(12)
    p = np.zeros((6000, 4))
    p[:, :3] = rng.integers(-2 ** 12, 2 ** 12, size=(6000, 3)) / 2.0 ** 12
    p[:, 3] = 1.0 / 6000
    r = oracle.reference_pipeline(p.astype(np.float32), 16, 64, 0.5, 0.01)
    d = oracle.reference_pipeline_d(p, 16, 64, 0.5, 0.01)
    assert np.array_equal(d["box"], r["box"].astype(np.float64))
    assert np.array_equal(d["tree"].cells, r["tree"].cells)
    assert np.array_equal(d["tree"].order, r["tree"].order)
    assert np.array_equal(d["sorted_keys"], r["sorted_keys"])
    assert np.array_equal(d["groups"], r["groups"])
    a = d["acc"][:, :3]
    assert np.sqrt(((r["acc"][:, :3] - a) ** 2).sum() / (a ** 2).sum()) < 1e-5


def test_fp64_oracle_multipoles_closed_forms_and_parallel_axis(oracle, bt):
    a = 0.25
    p = np.array([[a, 0, 0, 1], [-a, 0, 0, 1]], np.float64)
    cells = np.array([[1 << 27, 0xFFFFFFFF, 0, 2]], np.uint32)
    size, mono, q0, q1, ls = oracle.multipoles_d(cells, p, 0.5)
    assert mono[0, 3] == 2 and not mono[0, :3].any()
    assert q0[0, 0] == a * a and not q0[0, 1:].any() and not q1.any()
    assert ls[0, 0] == 2 * a and size[0, 3] == (2 * a / 0.5) ** 2
    pos, _ = bt.plummer(5000, seed=4)
    pos = pos.astype(np.float64)
    pos[:, :3] += np.random.default_rng(1).uniform(-1e-9, 1e-9, (5000, 3))  # digits below FP32
    t = oracle.build_tree(pos, 16)
    pt = pos[t.order]
    size, mono, q0, q1, ls = oracle.multipoles_d(t.cells, pt, 0.5)
    Q = np.zeros((len(t.cells), 3, 3))
    Q[:, 0, 0], Q[:, 1, 1], Q[:, 2, 2] = q0[:, 0], q0[:, 1], q0[:, 2]
    Q[:, 0, 1] = Q[:, 1, 0] = q0[:, 3]
    Q[:, 0, 2] = Q[:, 2, 0] = q1[:, 0]
    Q[:, 1, 2] = Q[:, 2, 1] = q1[:, 1]
    cells = t.cells.astype(np.int64)
    for c in np.where(cells[:, 1] != 0xFFFFFFFF)[0][:200]:
        f, n = cells[c, 1] & 0x1FFFFFFF, (cells[c, 1] >> 29) + 1
        kids = np.arange(f, f + n)
        m = mono[kids, 3]
        x = mono[kids, :3]
        M = m.sum()
        com = (m[:, None] * x).sum(0) / M
        dd = x - com
        Qp = ((m[:, None, None] * (Q[kids] + dd[:, :, None] * dd[:, None, :])).sum(0)) / M
        assert np.allclose(com, mono[c, :3], rtol=0, atol=1e-14 * max(1.0, np.abs(com).max()))
        assert np.allclose(Qp, Q[c], rtol=0, atol=1e-12 * ls[c, 0] ** 2)


def test_fp64_oracle_walk_theta_to_zero_is_direct_sum(oracle, bt):
    pos, _ = bt.plummer(2000, seed=2)
    pos = pos.astype(np.float64)
    r = oracle.reference_pipeline_d(pos, 16, 64, 1e-4, 0.01)
    assert r["stats"][2] == 0
    d = oracle.direct_fp64(r["pos_ph"], np.arange(2000, dtype=np.int32), 0.01)
    rel = np.linalg.norm(r["acc"][:, :3] - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
    assert rel.max() < 1e-12
    assert np.allclose(r["acc"][:, 3], d[:, 3], rtol=1e-12, atol=0)
    # the FP64 direct sum itself against numpy brute force
    x = r["pos_ph"][:64]
    dx = x[None, :, :3] - x[:, None, :3]
    r2 = (dx ** 2).sum(-1) + 0.01 ** 2
    acc = (x[None, :, 3, None] * dx / r2[..., None] ** 1.5).sum(1)
    d64 = oracle.direct_fp64(x, np.arange(64, dtype=np.int32), 0.01)
    assert np.allclose(d64[:, :3], acc, rtol=1e-13, atol=1e-13 * np.abs(acc).max())

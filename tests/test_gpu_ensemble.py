"""GPU tests of the batched paths: a forest of independent systems in one
context (one persistent walk over all systems' groups) against the oracle run
on each system alone (PH groups and leaf groups), the HIP-graph step against
the eager step (fixed angle, and learned angle with its per-step check), steps
with makeGroups' work on a second stream against steps in sequence, and the
autotuner.

Bars: per system, forces within 1e-5 relative RMS of the oracle walk on that
system alone and interaction counts equal; graph replays bit-identical to the
eager step after every step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THETA, EPS = 0.5, 0.01


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_forest_matches_each_system_alone(bt, oracle):
    pos, vel, off = bt.ensemble(nsys=6, seed=11, log2_min=10, log2_max=14)
    n = int(off[-1])
    t = bt.Treecode(EPS, THETA, 64)
    t.alloc(n, off)
    t.ptcl_h2d(pos, vel)
    t.buildTree(16)
    t.computeMultipoles()
    t.makeGroups(5, 64)
    t.computeForces()
    acc = t.acc(bt.ORDER_INPUT)
    cnt = t.counts()
    t.close()
    ncells = 0
    for k in range(6):
        pk = pos[off[k]:off[k + 1]]
        ref = oracle.reference_pipeline(pk, 16, 64, THETA, EPS)
        ncells += ref["tree"].ncells
        # oracle acc is in PH order of the system: back to input order
        order = ref["tree"].order[ref["perm"]]
        a_ref = np.empty_like(ref["acc"])
        a_ref[order] = ref["acc"]
        assert rel_rms(acc[off[k]:off[k + 1], :3], a_ref[:, :3]) <= 1e-5, k
    assert cnt["nsys"] == 6 and cnt["ncells"] == ncells


def test_graph_step_matches_eager_step(bt):
    pos, vel, off = bt.ensemble(nsys=5, seed=12, log2_min=11, log2_max=13)
    n = int(off[-1])
    ts = []
    for _ in range(2):
        t = bt.Treecode(EPS, THETA, 64)
        t.alloc(n, off)
        t.ptcl_h2d(pos, vel)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        ts.append(t)
    eager, graph = ts
    try:
        for step in range(4):
            eager.step(1.0 / 128, 16, 5, 64)
            graph.step_graph(1.0 / 128, 16, 5, 64, nsteps=1)
            pe, ve = eager.ptcl_d2h(bt.ORDER_INPUT)
            pg, vg = graph.ptcl_d2h(bt.ORDER_INPUT)
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), step
            assert np.array_equal(eager.acc(bt.ORDER_INPUT), graph.acc(bt.ORDER_INPUT)), step
        # several replays in one call equal as many eager steps
        for _ in range(3):
            eager.step(1.0 / 128, 16, 5, 64)
        graph.step_graph(1.0 / 128, 16, 5, 64, nsteps=3)
        assert np.array_equal(eager.ptcl_d2h(bt.ORDER_INPUT)[0], graph.ptcl_d2h(bt.ORDER_INPUT)[0])
    finally:
        eager.close()
        graph.close()


@pytest.mark.parametrize("nsys", [1, 5])
def test_overlapped_step_matches_serial_step(bt, nsys):
    """bt_set_overlap: the groups branch on a second stream gives bitwise the
    results of the phases in sequence, eager and graph-replayed."""
    if nsys == 1:
        pos, vel = bt.plummer(1 << 15, seed=13)
        off = None
        n = len(pos)
    else:
        pos, vel, off = bt.ensemble(nsys=nsys, seed=13, log2_min=11, log2_max=13)
        n = int(off[-1])
    ts = []
    for overlap in (True, False, True, False):
        t = bt.Treecode(EPS, THETA, 64)
        t.set_overlap(overlap)
        if off is None:
            t.alloc(n)
        else:
            t.alloc(n, off)
        t.ptcl_h2d(pos, vel)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        ts.append(t)
    try:
        for step in range(3):
            ts[0].step(1.0 / 128, 16, 5, 64)
            ts[1].step(1.0 / 128, 16, 5, 64)
            ts[2].step_graph(1.0 / 128, 16, 5, 64, nsteps=1)
            ts[3].step_graph(1.0 / 128, 16, 5, 64, nsteps=1)
            p0, v0 = ts[0].ptcl_d2h(bt.ORDER_INPUT)
            a0 = ts[0].acc(bt.ORDER_INPUT)
            g0 = ts[0].group_counts()
            for k in (1, 2, 3):
                pk, vk = ts[k].ptcl_d2h(bt.ORDER_INPUT)
                assert np.array_equal(p0, pk) and np.array_equal(v0, vk), (step, k)
                assert np.array_equal(a0, ts[k].acc(bt.ORDER_INPUT)), (step, k)
                assert np.array_equal(g0, ts[k].group_counts()), (step, k)
        tm = ts[1].timings()
        assert all(tm[k] > 0 for k in ("sort", "levels", "multipoles", "ph_sort", "groups", "walk"))
    finally:
        for t in ts:
            t.close()


def test_autotune_picks_a_candidate_and_stays_correct(bt, oracle):
    n = 30000
    pos, vel = bt.plummer(n, seed=13)
    t = bt.Treecode(EPS, THETA, 64)
    t.alloc(n)
    t.ptcl_h2d(pos, vel)
    best, rows = t.autotune(5)
    assert best["nleaf"] in (16, 32, 64) and best["ncrit"] in (32, 48, 64)
    assert best["blocks_per_cu"] in (4, 8, 12, 16) and best["stack_ints"] >= 1024
    assert len(rows) == 13 and all(r["walk_ms"] > 0 for r in rows)
    assert best["stack_ints"] >= 4 * max(r["deepest_stack"] for r in rows[9:])
    # the tuned set-up still computes the reference forces
    t.buildTree(best["nleaf"])
    t.computeMultipoles()
    t.makeGroups(5, best["ncrit"])
    t.computeForces()
    acc = t.acc(bt.ORDER_PH)
    ref = oracle.reference_pipeline(t.ptcl_d2h(bt.ORDER_INPUT)[0], best["nleaf"], best["ncrit"],
                                    THETA, EPS)
    t.close()
    assert rel_rms(acc[:, :3], ref["acc"][:, :3]) <= 1e-5


def test_forest_leaf_groups_match_each_system_alone(bt, oracle):
    """Leaf groups in a forest: every group lies in one system, and each
    system's forces equal the oracle's walk over that system's leaves."""
    pos, vel, off = bt.ensemble(nsys=4, seed=17, log2_min=10, log2_max=13)
    n = int(off[-1])
    t = bt.Treecode(EPS, THETA, 64)
    try:
        t.alloc(n, off)
        t.ptcl_h2d(pos, vel)
        t.set_group_mode(1)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        acc = t.acc(bt.ORDER_INPUT)
        ids = t.ids(bt.ORDER_PH)
        g = t.groups()
    finally:
        t.close()
    # a group's particles all belong to one system
    sys_of = np.searchsorted(off, ids, side="right") - 1
    for b, m in g:
        assert (sys_of[b:b + m] == sys_of[b]).all()
    for k in range(4):
        pk = pos[off[k]:off[k + 1]]
        ref = oracle.reference_pipeline(pk, 16, 64, THETA, EPS)
        lg = oracle.leaf_groups(ref["tree"].cells)
        a_o, _, _ = oracle.walk(lg, ref["tree"].level_range, ref["pos_tree"], ref["pos_tree"],
                                ref["tree"].cells, ref["size"], ref["mono"], ref["q0"], ref["q1"],
                                EPS)
        a_ref = np.empty_like(a_o)
        a_ref[ref["tree"].order] = a_o
        assert rel_rms(acc[off[k]:off[k + 1], :3], a_ref[:, :3]) <= 1e-5, k


def test_graph_step_with_learned_check_in_a_forest(bt):
    """A forest with the learned angle and its per-step check (the FP64
    direct sum per system): graph replays equal eager steps bitwise."""
    from bonsai2_amd import learn
    params, budget, lo, hi = learn.load()
    pos, vel, off = bt.ensemble(nsys=5, seed=19, log2_min=11, log2_max=13)
    n = int(off[-1])
    ts = []
    for _ in range(2):
        t = bt.Treecode(EPS, THETA, 64)
        t.alloc(n, off)
        t.ptcl_h2d(pos, vel)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.set_mlp_params(params)
        t.set_opening(1, budget, lo, hi)
        t.set_opening_check(128, 1e-3, 3)
        t.computeForces()
        ts.append(t)
    eager, graph = ts
    try:
        for step in range(3):
            eager.step(1.0 / 128, 16, 5, 64)
            graph.step_graph(1.0 / 128, 16, 5, 64, nsteps=1)
            assert np.array_equal(eager.ptcl_d2h(bt.ORDER_INPUT)[0],
                                  graph.ptcl_d2h(bt.ORDER_INPUT)[0]), step
            assert np.array_equal(eager.acc(bt.ORDER_INPUT), graph.acc(bt.ORDER_INPUT)), step
            assert eager.opening_check() == graph.opening_check(), step
        assert eager.opening_check()["checks"] == 4
    finally:
        eager.close()
        graph.close()


def test_full_ensemble_matches_systems_alone(bt):
    """The 256-system ensemble at full size (N_k = 2^U[10, 16], 3.87 M
    particles, theta 0.5, eps 0.01) in one context: the smallest, a middle
    and the largest system equal the same system run alone in a context of
    its own: the same groups with the same per-group interaction counts, and
    forces within 1e-6 relative RMS."""
    pos, vel, off = bt.ensemble(256, seed=19810614)
    n = int(off[-1])
    sizes = np.diff(off)
    t = bt.Treecode(EPS, THETA, 64)
    try:
        t.alloc(n, off)
        t.ptcl_h2d(pos, vel)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        acc = t.acc(bt.ORDER_INPUT)
        cnt = t.counts()
        groups = t.groups()
        gcnt = t.group_counts()
    finally:
        t.close()
    assert cnt["nsys"] == 256 and n > 3_000_000
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[128]), int(order[-1])]
    for k in picks:
        s = bt.Treecode(EPS, THETA, 64)
        try:
            s.alloc(int(sizes[k]))
            s.ptcl_h2d(pos[off[k]:off[k + 1]], vel[off[k]:off[k + 1]])
            s.buildTree(16)
            s.computeMultipoles()
            s.makeGroups(5, 64)
            s.computeForces()
            a_k = s.acc(bt.ORDER_INPUT)
            g_k = s.groups()
            c_k = s.group_counts()
        finally:
            s.close()
        mine = (groups[:, 0] >= off[k]) & (groups[:, 0] < off[k + 1])
        assert np.array_equal(groups[mine, 0] - off[k], g_k[:, 0]), k
        assert np.array_equal(groups[mine, 1], g_k[:, 1]), k
        assert np.array_equal(gcnt[mine], c_k), k
        print("system %d (n=%d): rel RMS %.2e" % (k, sizes[k],
                                                 rel_rms(acc[off[k]:off[k + 1], :3], a_k[:, :3])))
        assert rel_rms(acc[off[k]:off[k + 1], :3], a_k[:, :3]) <= 1e-6, k

"""GPU tests of the failure modes the reference guards with device asserts or
not at all, each reached through the library's own bounds checks (every kernel
returns normally):

* walk-stack overflow -> BT_ERR_STACK (the reference returns a sentinel and
  then asserts, code/computeForces.cu:302-303, :565), eagerly, in a step and
  in a captured step replayed several times;
* cell capacity: an input needing more than the initial n + 1024 + 64 cells
  grows the arrays once, in buildTree, in a step and in a captured step (the
  reference asserts, code/buildTree.cu:412, :422-429), and the tree still
  equals the oracle's;
* more than NLEAF coincident particles -> BT_FLAG_DEPTH_CAP (the reference has
  no depth guard) with tree and forces equal to the oracle's;
* particles on and beyond the quantised root box's edge
  (code/buildTree.cu:147-159): box, tree, keys and groups bit-exact, forces
  within 1e-5 of the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BT_ERR_STACK = -5


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def forces_of(bt, pos, vel, eps=0.01, theta=0.4, nleaf=16, ncrit=64):
    t = bt.Treecode(eps, theta, ncrit)
    t.alloc(len(pos))
    t.ptcl_h2d(pos, vel)
    t.buildTree(nleaf)
    t.computeMultipoles()
    t.makeGroups(5, ncrit)
    t.computeForces()
    return t


def clumps(nclump=1024, seed=5):
    """nclump clumps of NLEAF + 1 = 17 particles, 9 and 8 on either side of a
    level-21 cell boundary 2^-23 away: each clump is a chain of single-child
    nodes down to level 21 (about 1.08 cells per particle), plus two anchor
    particles that fix the root box to [-1, 1]^3."""
    rng = np.random.default_rng(seed)
    k = rng.integers(-2 ** 18, 2 ** 18, size=(nclump, 3))
    cen = (k * 2 + 1).astype(np.float64) / 2 ** 20 * 0.9
    cen = (np.round(cen * 2 ** 20 / 2) * 2 + 1) / 2 ** 20      # odd multiples of 2^-20
    off = np.zeros((17, 3))
    off[:9, 0] = -2.0 ** -23
    off[9:, 0] = 2.0 ** -23
    p = (cen[:, None, :] + off[None]).reshape(-1, 3)
    p = np.concatenate([p, [[0.9375] * 3, [-0.9375] * 3]])
    pos = np.zeros((len(p), 4), np.float32)
    pos[:, :3] = p
    pos[:, 3] = 1.0 / len(p)
    vel = np.zeros_like(pos)
    vel[:, 3] = pos[:, 3]
    return pos, vel


def drifting_into_clumps(dt):
    """Spread-out particles whose velocities take them onto the clumps in one
    drift of dt (the half kick before it moves them by ~|a| dt^2 / 2, far
    below the clumps' 2^-23 structure at dt = 2^-16)."""
    tgt, _ = clumps()
    n = len(tgt)
    rng = np.random.default_rng(9)
    x0 = tgt.copy()
    x0[:-2, :3] += rng.uniform(-1e-3, 1e-3, (n - 2, 3)).astype(np.float32)
    v = np.zeros_like(x0)
    v[:, :3] = ((tgt[:, :3].astype(np.float64) - x0[:, :3]) / dt).astype(np.float32)
    v[:, 3] = x0[:, 3]
    return x0, v


def initial_cell_capacity(n, nsys=1):
    return n + 1024 * nsys + 64


# ---- walk stack ------------------------------------------------------------

def test_walk_stack_overflow_is_an_error(bt):
    """The fixed-angle walk keeps up to 512 stack entries in LDS here and
    spills the rest to a 64-entry global stack: a tree whose walk needs more
    than 576 entries overflows, in a walk, an eager step and a captured step."""
    n = 1 << 20
    pos, vel = bt.clustered(n, seed=22)
    t = forces_of(bt, pos, vel, eps=0.002)
    try:
        ref = t.acc(bt.ORDER_PH)
        deepest = t.walk_stack_max()
        print("deepest stack %d entries; walking with 512 in LDS + 64" % deepest, flush=True)
        assert deepest > 512 + 64
        t.set_walk_lds_stack(512)
        t.set_walk_options(kahan=1, stack_ints=64)
        with pytest.raises(bt.BonsaiError, match=r"\(%d\).*stack" % BT_ERR_STACK):
            t.computeForces()
        # the same context walks correctly again with a stack that fits
        t.set_walk_options(kahan=1, stack_ints=16384)
        t.computeForces()
        assert np.array_equal(t.acc(bt.ORDER_PH), ref)
        # an eager step: the error, and the particles hold the step's first
        # half kick and drift only
        p0, v0 = t.ptcl_d2h(bt.ORDER_INPUT)
        t.set_walk_options(kahan=1, stack_ints=64)
        with pytest.raises(bt.BonsaiError, match="stack"):
            t.step(1.0 / 128, 16, 5, 64)
        p1, _ = t.ptcl_d2h(bt.ORDER_INPUT)
        assert not np.array_equal(p0, p1) and np.isfinite(p1).all()
        # a captured step replayed three times: the error of the first replay
        # is kept to the end (sticky), and the later replays move nothing
        t.set_walk_options(kahan=1, stack_ints=16384)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        t.set_walk_options(kahan=1, stack_ints=64)
        with pytest.raises(bt.BonsaiError, match="replay 1 of 3"):
            t.step_graph(1.0 / 128, 16, 5, 64, nsteps=3)
        t.set_walk_options(kahan=1, stack_ints=16384)
        t.buildTree(16)
        t.computeMultipoles()
        t.makeGroups(5, 64)
        t.computeForces()
        t.step_graph(1.0 / 128, 16, 5, 64, nsteps=2)
        assert np.isfinite(t.acc(bt.ORDER_PH)).all()
    finally:
        t.close()


def test_learned_walk_stack_overflow_is_an_error(bt):
    """The learned walk keeps up to 512 stack entries in LDS here and spills
    the rest to a 64-entry global stack: a tree whose walk needs more than
    576 entries overflows."""
    n = 1 << 20
    pos, vel = bt.clustered(n, seed=22)
    t = forces_of(bt, pos, vel, eps=0.002)
    try:
        t.use_learned_theta(check_samples=0)   # no per-step check: a deterministic walk
        t.computeForces()
        ref = t.acc(bt.ORDER_PH)
        deepest = t.walk_stack_max()
        print("learned walk: deepest stack %d entries" % deepest, flush=True)
        assert deepest > 512 + 64
        t.set_walk_options(kahan=1, stack_ints=64)
        t.set_walk_lds_stack(512)
        with pytest.raises(bt.BonsaiError, match="stack"):
            t.computeForces()
        t.set_walk_options(kahan=1, stack_ints=16384)
        t.set_walk_lds_stack(1024)
        t.computeForces()
        assert np.array_equal(t.acc(bt.ORDER_PH), ref)
    finally:
        t.close()


# ---- cell capacity -----------------------------------------------------------

def test_cell_capacity_grows_in_build(bt, oracle):
    pos, vel = clumps()
    n = len(pos)
    ref = oracle.reference_pipeline(pos, 16, 64, 0.4, 0.01)
    assert ref["tree"].ncells > initial_cell_capacity(n)
    assert ref["tree"].nlevels == 21 and ref["tree"].capped == 0
    print("clumps: n=%d cells=%d" % (n, ref["tree"].ncells), flush=True)
    t = forces_of(bt, pos, vel)
    try:
        c = t.counts()
        assert c["ncells"] == ref["tree"].ncells and c["flags"] == 0
        assert np.array_equal(t.tree(multipoles=False)["cells"], ref["tree"].cells)
        assert np.array_equal(t.groups(), ref["groups"])
        acc = t.acc(bt.ORDER_PH)
        assert np.isfinite(acc).all()
        assert rel_rms(acc[:, :3], ref["acc"][:, :3]) <= 1e-5
    finally:
        t.close()


def test_cell_capacity_grows_in_a_step_and_in_a_captured_step(bt, oracle):
    dt = 2.0 ** -16
    x0, v0 = drifting_into_clumps(dt)
    n = len(x0)
    assert oracle.build_tree(x0, 16).ncells < initial_cell_capacity(n)
    eager = forces_of(bt, x0, v0)
    graph = forces_of(bt, x0, v0)
    try:
        assert eager.counts()["ncells"] < initial_cell_capacity(n)
        eager.step(dt, 16, 5, 64)
        # after one step the particles sit on the clumps: their tree needed
        # more cells than the arrays first held, and equals the oracle's
        pt, _ = eager.ptcl_d2h(bt.ORDER_TREE)
        ref = oracle.build_tree(pt, 16)
        print("step 1: cells %d (first capacity %d)" % (ref.ncells, initial_cell_capacity(n)),
              flush=True)
        assert ref.ncells > initial_cell_capacity(n)
        assert eager.counts()["ncells"] == ref.ncells
        assert np.array_equal(eager.tree(multipoles=False)["cells"], ref.cells)
        assert np.isfinite(eager.acc(bt.ORDER_PH)).all()
        eager.step(dt, 16, 5, 64)
        # the captured step hits the same overflow in its first replay: the
        # cells grow, that step is finished and the second runs eagerly
        graph.step_graph(dt, 16, 5, 64, nsteps=2)
        pe, ve = eager.ptcl_d2h(bt.ORDER_INPUT)
        pg, vg = graph.ptcl_d2h(bt.ORDER_INPUT)
        assert np.array_equal(pe, pg) and np.array_equal(ve, vg)
    finally:
        eager.close()
        graph.close()


# ---- coincident particles and the box edge ------------------------------------

def test_coincident_particles_set_depth_cap(bt, oracle):
    n0 = 4096
    pos, vel = bt.plummer(n0, seed=23)
    extra = np.repeat(pos[:1], 40, axis=0)    # 41 particles at one point (> NLEAF)
    pos = np.concatenate([pos, extra])
    vel = np.concatenate([vel, np.repeat(vel[:1], 40, axis=0)])
    pos[:, 3] = vel[:, 3] = 1.0 / len(pos)
    ref = oracle.reference_pipeline(pos, 16, 64, 0.4, 0.01)
    assert ref["tree"].capped == 1
    t = forces_of(bt, pos, vel)
    try:
        c = t.counts()
        assert c["flags"] & 1                      # BT_FLAG_DEPTH_CAP
        assert c["nlevels"] == ref["tree"].nlevels == 21
        assert np.array_equal(t.tree(multipoles=False)["cells"], ref["tree"].cells)
        assert np.array_equal(t.groups(), ref["groups"])
        acc = t.acc(bt.ORDER_PH)
        assert rel_rms(acc[:, :3], ref["acc"][:, :3]) <= 1e-5
        assert rel_rms(acc[:, 3], ref["acc"][:, 3]) <= 1e-5
    finally:
        t.close()


@pytest.mark.parametrize("shift", [0.0, 0.7])
def test_particles_on_the_quantised_box_edge(bt, oracle, shift):
    """shift 0: extremes at exactly -1 and +1, so hsize = 1 equals the half
    extent and the top particle sits on the box's upper face (its PH
    coordinate is 2^21, one past the last cell).  shift 0.7: the centre
    0.7 is quantised down to a multiple of 2^-20, leaving the top particle
    just outside the box."""
    n = 8192
    pos, vel = bt.plummer(n, seed=24)
    p = pos[:, :3].astype(np.float64)
    p = p / np.abs(p).max() * 0.999
    p[0] = [1.0, 1.0, 1.0]
    p[1] = [-1.0, -1.0, -1.0]
    pos[:, :3] = (p + shift).astype(np.float32)
    ref = oracle.reference_pipeline(pos, 16, 64, 0.5, 0.01)
    box = ref["box"]
    assert box[3] == 1.0
    assert pos[:, 0].max() >= box[0] + box[3]
    t = forces_of(bt, pos, vel, theta=0.5)
    try:
        assert np.array_equal(t.boxes()[0], box)
        assert np.array_equal(t.tree(multipoles=False)["cells"], ref["tree"].cells)
        keys, perm = t.ph()
        assert np.array_equal(keys, ref["sorted_keys"]) and np.array_equal(perm, ref["perm"])
        assert np.array_equal(t.groups(), ref["groups"])
        acc = t.acc(bt.ORDER_PH)
        assert np.isfinite(acc).all()
        assert rel_rms(acc[:, :3], ref["acc"][:, :3]) <= 1e-5
    finally:
        t.close()

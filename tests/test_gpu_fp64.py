"""GPU parity of Treecode<double> (code/Particle4.h:106-125; the reference
instantiates float only, code/TreecodeInstances.h:3) against the oracle's FP64
pipeline.

The inputs are the float32 samples made float64 and moved by a seeded offset
of ~1e-9, so that the FP64 digits matter.  Bars: root box, tree, order, PH
keys, permutation and groups bit-exact; moments within 1e-12 (relative to
l^2 for Q) of the oracle's two-pass FP64 sums; every group's interaction
counts exact against the oracle's FP64 walk over the GPU's own tree and
moments, forces within 1e-11 relative RMS of it (different FP64 summation
orders and the P2C rearranged on the 3Q record), and within 1e-6 of the
oracle's walk on its own moments (an opening decision exactly on the boundary
may flip); FP64 direct sum and energies within 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (name, n, theta, eps, nleaf, ncrit, level_split, ic)
    ("plummer_2p16_theta0.4", 1 << 16, 0.4, 0.01, 16, 64, 5, "plummer"),
    ("clustered_small", 1 << 15, 0.5, 0.002, 16, 64, 5, "clustered"),
]


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def fp64_input(bt, ic, n, seed=19810614):
    pos, vel = getattr(bt, ic)(n, seed=seed)
    rng = np.random.default_rng(seed + 64)
    p = pos.astype(np.float64)
    v = vel.astype(np.float64)
    p[:, :3] += rng.uniform(-1e-9, 1e-9, (n, 3))
    v[:, :3] += rng.uniform(-1e-9, 1e-9, (n, 3))
    return p, v


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def run(request, bt, oracle):
    name, n, theta, eps, nleaf, ncrit, ls, ic = request.param
    print("case", name, flush=True)
    pos, vel = fp64_input(bt, ic, n)
    t = bt.Treecode(eps, theta, ncrit, fp64=True)
    t.alloc(n)
    t.ptcl_h2d(pos, vel)
    t.buildTree(nleaf)
    t.computeMultipoles()
    t.makeGroups(ls, ncrit)
    stats = t.computeForces()
    ref = oracle.reference_pipeline_d(pos, nleaf, ncrit, theta, eps, level_split=ls)
    yield dict(t=t, pos=pos, vel=vel, ref=ref, stats=stats, theta=theta, eps=eps, n=n)
    t.close()


def test_box_tree_keys_groups_bit_exact(run):
    t, ref = run["t"], run["ref"]
    assert t.boxes().dtype == np.float64
    assert np.array_equal(t.boxes()[0], ref["box"])
    tr = t.tree(multipoles=False)
    assert t.counts()["ncells"] == ref["tree"].ncells
    assert np.array_equal(tr["cells"], ref["tree"].cells)
    assert np.array_equal(tr["level_range"], ref["tree"].level_range)
    assert np.array_equal(t.ids(1), ref["tree"].order)
    keys, perm = t.ph()
    assert np.array_equal(keys, ref["sorted_keys"]) and np.array_equal(perm, ref["perm"])
    assert np.array_equal(t.groups(), ref["groups"])
    # the particles come back in FP64, bit for bit
    p_in, v_in = t.ptcl_d2h(0)
    assert np.array_equal(p_in, run["pos"]) and np.array_equal(v_in, run["vel"])


def test_moments_fp64(run):
    t, ref = run["t"], run["ref"]
    tr = t.tree()
    l = ref["ls"][:, 0]
    assert np.allclose(tr["mono"][:, 3], ref["mono"][:, 3], rtol=1e-13, atol=0)
    com_err = np.abs(tr["mono"][:, :3] - ref["mono"][:, :3]).max(axis=1)
    assert (com_err <= 1e-12 * l + 4 * np.spacing(np.abs(ref["mono"][:, :3]).max(axis=1))).all()
    q = np.concatenate([tr["q0"], tr["q1"]], 1)
    qr = np.concatenate([ref["q0"], ref["q1"]], 1)
    assert (np.abs(q - qr).max(axis=1) <= 1e-12 * l * l + 1e-300).all()
    # cellOp = l/theta + s; s = |bbox centre - com| carries a few ulp(|com|)
    op_g = np.sqrt(tr["size"][:, 3])
    op_r = np.sqrt(ref["size"][:, 3])
    ulp = np.spacing(np.abs(ref["mono"][:, :3]).max(axis=1))
    assert (np.abs(op_g - op_r) <= 1e-12 * op_r + 8 * ulp).all()


def test_walk_counts_exact_and_forces(run, oracle):
    t, ref, n = run["t"], run["ref"], run["n"]
    tr = t.tree()
    pos_tree = t.ptcl_d2h(1)[0]
    pos_ph = t.ptcl_d2h(2)[0]
    acc_o, st_o, pot_o, cnt_o = oracle.walk_d(t.groups(), tr["level_range"], pos_ph, pos_tree,
                                              tr["cells"], tr["size"], tr["mono"], tr["q0"],
                                              tr["q1"], run["eps"])
    assert np.array_equal(t.group_counts(), cnt_o)
    st = run["stats"]
    assert tuple(st) == (st_o[0] / n, float(st_o[1]), st_o[2] / n, float(st_o[3]))
    acc = t.acc(2)
    assert acc.dtype == np.float64
    rms = rel_rms(acc[:, :3], acc_o[:, :3])
    print("FP64 walk vs oracle on the GPU's moments: rel RMS %.3e" % rms)
    assert rms <= 1e-11
    assert rel_rms(acc[:, 3], acc_o[:, 3]) <= 1e-11
    assert abs(t.grav_potential - pot_o) <= 1e-11 * abs(pot_o)
    # against the oracle's own FP64 pipeline (a decision exactly on the
    # boundary may flip where the moments differ in the last bit)
    assert rel_rms(acc[:, :3], ref["acc"][:, :3]) <= 1e-6


def test_monopole_walk_fp64(run, oracle):
    t, n = run["t"], run["n"]
    t.set_quadrupole(False)
    try:
        st = t.computeForces()
        acc = t.acc(2)
        cnt = t.group_counts()
    finally:
        t.set_quadrupole(True)
        t.computeForces()
    tr = t.tree()
    acc_o, st_o, _, cnt_o = oracle.walk_d(t.groups(), tr["level_range"], t.ptcl_d2h(2)[0],
                                          t.ptcl_d2h(1)[0], tr["cells"], tr["size"], tr["mono"],
                                          tr["q0"], tr["q1"], run["eps"], monopole=True)
    assert np.array_equal(cnt, cnt_o)
    assert tuple(st) == (st_o[0] / n, float(st_o[1]), st_o[2] / n, float(st_o[3]))
    assert rel_rms(acc[:, :3], acc_o[:, :3]) <= 1e-11


def test_direct_sum_and_energies_fp64(run, oracle):
    t, ref = run["t"], run["ref"]
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(run["n"], size=256, replace=False)).astype(np.int32)
    gd = t.direct_fp64(idx)
    od = oracle.direct_fp64(ref["pos_ph"], idx, run["eps"])
    assert np.allclose(gd, od, rtol=1e-12, atol=1e-13 * np.abs(od).max())
    a = t.acc(2)[idx, :3]
    e = np.linalg.norm(a - gd[:, :3], axis=1) / np.linalg.norm(gd[:, :3], axis=1)
    p50, p99 = np.percentile(e, [50, 99])
    print("FP64 walk vs FP64 direct sum: p50 %.3e p99 %.3e" % (p50, p99))
    assert p99 < 2e-2
    ek, ep = t.computeEnergies()
    pos, vel = t.ptcl_d2h(1)
    k = 0.5 * (pos[:, 3] * (vel[:, :3] ** 2).sum(1)).sum()
    acc_tree = t.acc(1)
    w = 0.5 * (pos[:, 3] * (acc_tree[:, 3] + pos[:, 3] / run["eps"])).sum()
    assert abs(ek[0] - k) <= 1e-12 * abs(k) and abs(ep[0] - w) <= 1e-12 * abs(w)


def test_fp64_step_tracks_fp32_step(run, bt):
    """One kick-drift-kick step in FP64 and in FP32 from the same particles:
    the states agree to FP32 accuracy, and a float entry point on the FP64
    context is an error, not a misread."""
    t, n = run["t"], run["n"]
    pos, vel = run["pos"], run["vel"]
    s = bt.Treecode(run["eps"], run["theta"], 64)
    try:
        s.alloc(n)
        s.ptcl_h2d(pos.astype(np.float32), vel.astype(np.float32))
        for tc in (t, s):
            tc.buildTree(16)
            tc.computeMultipoles()
            tc.makeGroups(5, 64)
            tc.computeForces()
            tc.step(1.0 / 64, 16, 5, 64)
        p64, v64 = t.ptcl_d2h(0)
        p32, v32 = s.ptcl_d2h(0)
        assert np.isfinite(p64).all()
        assert np.abs(p64[:, :3] - p32[:, :3]).max() < 1e-5
        assert rel_rms(v32[:, :3], v64[:, :3]) < 1e-5
        # a float data entry point on the FP64 context
        buf = np.zeros((n, 4), np.float32)
        assert t._L.bt_get_acc(t._h, buf.ctypes.data, 0) == -3    # BT_ERR_STATE
        assert b"_f64" in t._L.bt_last_error(t._h)
        assert not buf.any()
    finally:
        s.close()

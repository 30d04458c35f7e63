"""GPU checks at the benchmark sizes (N = 2^22 Plummer, N = 2^24 halo+disk,
N = 2^24 clustered), where the CPU oracle would take too long:
size-independent properties of every stage.  Tree: the invariants of
SWtest/treebuild.cu:1477 (leaf counts sum to N, leaves <= NLEAF, nodes >
NLEAF); PH keys sorted and the permutation a bijection; groups tile [0, N)
with <= NCRIT each; root moments equal the FP64 sums over all particles;
forces against the FP64 direct sum on seeded samples; two force evaluations
bit-identical; total momentum of the forces small; energy conserved over 8
leapfrog steps (Plummer and halo+disk); and, on the deep clustered tree,
compensated against plain FP32 accumulation over the same interaction
lists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # name, n, ic, theta, eps, nleaf, ncrit, p99 bound
    ("plummer_2p22", 1 << 22, "plummer", 0.5, 0.005, 16, 64, 1e-3),
    ("halo_disk_2p24", 1 << 24, "halo_disk", 0.4, 0.002, 16, 64, 1e-3),
    ("clustered_2p24", 1 << 24, "clustered", 0.5, 0.002, 16, 64, 1e-3),
]


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def p99_vs_direct(t, n, seed=3, ns=1024):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=ns, replace=False)).astype(np.int32)
    d = t.direct_fp64(idx)
    a = t.acc(2)[idx, :3].astype(np.float64)
    e = np.linalg.norm(a - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
    return np.percentile(e, [50, 99])


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def big(request, bt):
    name, n, ic, theta, eps, nleaf, ncrit, p99 = request.param
    print("case", name, flush=True)
    make = {"plummer": bt.plummer, "halo_disk": bt.halo_disk, "clustered": bt.clustered}[ic]
    pos, vel = make(n)
    t = bt.Treecode(eps, theta, ncrit)
    t.alloc(n)
    t.ptcl_h2d(pos, vel)
    t.buildTree(nleaf)
    t.computeMultipoles()
    t.makeGroups(5, ncrit)
    st = t.computeForces()
    yield dict(t=t, pos=pos, vel=vel, n=n, nleaf=nleaf, ncrit=ncrit, p99=p99, eps=eps, st=st,
               name=name)
    t.close()


def test_tree_invariants_full_size(big):
    t = big["t"]
    tr = t.tree(multipoles=True)
    cells = tr["cells"].astype(np.int64)
    leaf = cells[:, 1] == 0xFFFFFFFF
    np_ = cells[:, 3] - cells[:, 2]
    assert np_[leaf].sum() == big["n"]
    assert (np_[leaf] <= big["nleaf"]).all() or t.counts()["flags"] != 0
    assert (np_[~leaf] > big["nleaf"]).all()
    # root moments: the level-1 cells' masses and mass-weighted centres
    lr = tr["level_range"].reshape(-1, 2)
    l1 = slice(lr[1, 0], lr[1, 1] + 1)
    mono = tr["mono"][l1].astype(np.float64)
    p = big["pos"].astype(np.float64)
    M = p[:, 3].sum()
    assert abs(mono[:, 3].sum() - M) <= 1e-6 * M
    com = (mono[:, :3] * mono[:, 3:4]).sum(0) / mono[:, 3].sum()
    com_ref = (p[:, :3] * p[:, 3:4]).sum(0) / M
    assert np.abs(com - com_ref).max() <= 1e-5 * np.abs(p[:, :3]).max()


def test_ph_order_and_groups_full_size(big):
    t = big["t"]
    keys, perm = t.ph()
    assert (keys[1:] >= keys[:-1]).all()
    assert np.array_equal(np.sort(perm), np.arange(big["n"]))
    g = t.groups()
    assert g[0, 0] == 0 and (g[1:, 0] == g[:-1, 0] + g[:-1, 1]).all()
    assert g[-1, 0] + g[-1, 1] == big["n"]
    assert (g[:, 1] >= 1).all() and (g[:, 1] <= big["ncrit"]).all()


def test_forces_against_fp64_direct_full_size(big):
    p50, p99 = p99_vs_direct(big["t"], big["n"])
    print("%s p50 %.2e p99 %.2e" % (big["name"], p50, p99))
    assert p99 <= big["p99"]


def test_compensated_and_plain_sums_full_size(big):
    """The clustered configuration's comparison: plain FP32 accumulation
    (kahan = 0) walks the same lists (equal interaction statistics), its
    forces stay within 1e-5 relative RMS of the compensated ones, and both
    meet the p99 bound against the FP64 direct sum."""
    if big["name"] != "clustered_2p24":
        pytest.skip("compensated against plain sums: the clustered case")
    t = big["t"]
    st_k = t.computeForces()
    a_k = t.acc(2).copy()
    k50, k99 = p99_vs_direct(t, big["n"], seed=5)
    t.set_walk_options(kahan=0)
    try:
        st_p = t.computeForces()
        a_p = t.acc(2)
        p50, p99 = p99_vs_direct(t, big["n"], seed=5)
    finally:
        t.set_walk_options(kahan=1)
        t.computeForces()   # leave the compensated forces for the tests that follow
    d = rel_rms(a_p[:, :3], a_k[:, :3])
    print("compensated p50 %.4e p99 %.4e; plain p50 %.4e p99 %.4e; rel RMS difference %.2e; "
          "levels %d" % (k50, k99, p50, p99, d, t.counts()["nlevels"]))
    assert tuple(st_p) == tuple(st_k)
    assert 0.0 < d <= 1e-5
    assert k99 <= big["p99"] and p99 <= big["p99"]


def test_forces_deterministic_and_momentum_full_size(big):
    t = big["t"]
    a1 = t.acc(2)
    st = t.computeForces()
    assert tuple(st) == tuple(big["st"])
    a2 = t.acc(2)
    assert np.array_equal(a1, a2)
    pos, _ = t.ptcl_d2h(0)         # input order
    a_in = t.acc(0)[:, :3].astype(np.float64)
    m = pos[:, 3:4].astype(np.float64)
    net = np.abs((m * a_in).sum(0)).max()
    assert net <= 1e-3 * (m * np.abs(a_in)).sum(0).max()


def test_energy_conserved_over_steps(big):
    if big["name"] == "clustered_2p24":
        pytest.skip("the clustered configuration specifies no integration")
    t = big["t"]
    ek0, ep0 = t.computeEnergies()
    e0 = ek0.sum() + ep0.sum()
    for _ in range(8):
        t.step(1.0 / 128, big["nleaf"], 5, big["ncrit"])
    ek1, ep1 = t.computeEnergies()
    e1 = ek1.sum() + ep1.sum()
    print("dE/E %.2e" % (abs(e1 - e0) / abs(e0)))
    assert abs(e1 - e0) <= 1e-5 * abs(e0)

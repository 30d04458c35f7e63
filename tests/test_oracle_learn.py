"""Pins of the oracle's learned-opening-angle functions (CPU only): FP16
rounding, pair features, the 8-32-32-1 network and its gradient, the walk
with per-pair angles and the training-pair generator."""
import numpy as np
import pytest



@pytest.fixture(scope="module")
def plummer_small(oracle):
    return oracle.plummer_ref_f4(3000, seed=7, rmax=22.8)[0]


def rand_params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(1377).astype(np.float32) * np.float32(scale)
    return p


def test_f16_round_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(200000) * 10.0 ** rng.integers(-9, 6, 200000)).astype(np.float32)
    special = np.array([65504, 65519.99, 65520, 6.1e-5, 6.0e-8, 5.96e-8, 2.9802322e-08, 3e-8,
                        1e-10, 0.0, -0.0, np.inf, -np.inf, 1.00048828125, 1.000732421875],
                       np.float32)
    x = np.concatenate([x, special, -special])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(oracle.f16_round(x), ref)


def test_pair_features_closed_form(oracle):
    # group at the origin, half sizes (0.1, 0.2, 0.05); cell com at (3, 0, 4)
    pin = np.array([[0, 0, 0, 0.1, 0.2, 0.05, 48, 3, 0, 4, 0.25, 1.0,
                     0.01, 0.02, 0.03, 0.001, 0.002, 0.003, 8.0, 1.0]], np.float32)
    f = oracle.pair_features(pin)[0]
    f32 = np.float32
    d = f32(5.0)
    rg = np.sqrt(f32(0.1) * f32(0.1) + f32(0.2) * f32(0.2) + f32(0.05) * f32(0.05))
    assert f[0] == rg / (rg + d)
    assert f[1] == f32(0.75)
    assert f[2] == f32(1.0) / (f32(1.0) + d)
    assert f[3] == d / (d + f32(8.0))
    assert f[4] == f32(0.25)
    q = np.array([0.01, 0.02, 0.03, 0.001, 0.002, 0.003], np.float32)
    qn = np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + f32(2) * (q[3] * q[3] + q[4] * q[4]
                                                                      + q[5] * q[5]))
    assert f[5] == qn / (qn + f32(0.25) * f32(1) * f32(1) + f32(1e-30))
    assert f[6] == f32(0.8)          # |o_z| / d
    assert f[7] == f32(0.0)          # longest group axis is y, o_y = 0
    # coincident com: direction cosines 0, group-radius feature 1
    pin2 = pin.copy()
    pin2[0, 7:10] = 0
    f2 = oracle.pair_features(pin2)[0]
    assert f2[6] == 0 and f2[7] == 0 and abs(f2[0] - 1) < 1e-6


def test_mlp_forward_hand_case(oracle):
    """One path through the network: W1[0][2] = 1, W2[0][0] = 1, w3[0] = 1."""
    p = np.zeros(1377, np.float32)
    p[0 * 8 + 2] = 1.0          # W1[0][2]
    p[288 + 0] = 1.0            # W2[0][0]
    p[1344] = 1.0               # w3[0]
    p[1376] = 0.5               # b3
    x = np.zeros((1, 8), np.float32)
    x[0, 2] = 0.3
    silu = lambda z: z / (1 + np.exp(-z))
    z1 = float(np.float16(0.3))
    a1 = float(np.float16(np.float32(silu(z1))))
    y_ref = 0.5 + silu(a1) + 31 * 0.0  # other hidden units: silu(0) = 0
    assert abs(oracle.mlp_forward(p, x)[0] - y_ref) < 1e-6


def test_mlp_gradient_finite_differences(oracle):
    rng = np.random.default_rng(5)
    p = rand_params(3, 0.4)
    x = rng.random((64, 8)).astype(np.float32)
    t = rng.standard_normal(64).astype(np.float32)
    loss, g = oracle.mlp_grad(p, x, t, round16=False)
    idx = np.concatenate([[0, 7, 255, 256, 287, 288, 700, 1311, 1312, 1343, 1344, 1375, 1376],
                          rng.integers(0, 1377, 20)])
    for i in idx:
        h = 1e-2 * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((oracle.mlp_forward(pp, x, round16=False).astype(np.float64) - t) ** 2)
        lm = np.mean((oracle.mlp_forward(pm, x, round16=False).astype(np.float64) - t) ** 2)
        fd = (lp - lm) / (2 * h)
        # the forward pass returns FP32 outputs: ~1e-5 absolute noise in fd
        assert abs(fd - g[i]) <= 2e-3 * abs(fd) + 2e-5, (i, fd, g[i])
    assert abs(loss - np.mean((oracle.mlp_forward(p, x, round16=False) - t) ** 2)) < 1e-5


@pytest.fixture(scope="module")
def small_tree(oracle, plummer_small):
    pos = plummer_small
    ref = oracle.reference_pipeline(pos, 16, 64, 0.5, 0.01)
    return ref


def test_constant_network_walk_equals_fixed_theta(oracle, small_tree):
    """A network whose output is the constant log2(0.5) gives every pair the
    angle 0.5: the walk must equal the fixed-angle walk (inv_theta 2)."""
    r = small_tree
    p = np.zeros(1377, np.float32)
    p[1376] = -1.0
    ng = len(r["groups"])
    gscale = oracle.group_scale(r["groups"], r["box"][3], 1.0)
    acc_l, st_l, pot_l, cnt_l = oracle.walk_learned(
        r["groups"], r["tree"].level_range, r["pos_ph"], r["pos_tree"], r["tree"].cells,
        r["size"], r["mono"], r["q0"], r["q1"], r["ls"], 0.01, p, 0.0, 0.2, 1.2, gscale)
    acc_f, st_f, pot_f, cnt_f = oracle.walk(
        r["groups"], r["tree"].level_range, r["pos_ph"], r["pos_tree"], r["tree"].cells,
        r["size"], r["mono"], r["q0"], r["q1"], 0.01, cell_ls=r["ls"],
        group_inv_theta=np.full(ng, 2.0, np.float32), counts=True)
    assert np.array_equal(acc_l, acc_f) and np.array_equal(st_l, st_f)
    assert np.array_equal(cnt_l, cnt_f)
    # and the clamp: a huge output is limited to theta_hi
    p[1376] = 10.0
    _, st_hi, _, _ = oracle.walk_learned(
        r["groups"], r["tree"].level_range, r["pos_ph"], r["pos_tree"], r["tree"].cells,
        r["size"], r["mono"], r["q0"], r["q1"], r["ls"], 0.01, p, 0.0, 0.2, 0.8, gscale)
    _, st_08, _ = oracle.walk(
        r["groups"], r["tree"].level_range, r["pos_ph"], r["pos_tree"], r["tree"].cells,
        r["size"], r["mono"], r["q0"], r["q1"], 0.01, cell_ls=r["ls"],
        group_inv_theta=np.full(ng, np.float32(1) / np.float32(0.8), np.float32))
    assert np.array_equal(st_hi, st_08)


def test_learn_pairs_boundary_and_error(oracle, small_tree):
    r = small_tree
    pos_ph = r["pos_ph"]
    n = len(pos_ph)
    gsel = np.array([0, 5, len(r["groups"]) // 2, len(r["groups"]) - 1], np.int32)
    pidx = np.concatenate([np.arange(r["groups"][g][0], r["groups"][g][0] + r["groups"][g][1])
                           for g in gsel]).astype(np.int32)
    d = oracle.direct_fp64(pos_ph, pidx, 0.01)
    amag = np.ones(n)
    amag[pidx] = np.linalg.norm(d[:, :3], axis=1)
    gscale = oracle.group_scale(r["groups"], r["box"][3], 1.0)
    out = oracle.learn_pairs(gsel, r["groups"], r["tree"].level_range, pos_ph, r["pos_tree"],
                             r["tree"].cells, r["size"], r["mono"], r["q0"], r["q1"], r["ls"],
                             gscale, 0.25, 1.0, 0.01, amag)
    assert out["n"] > 50
    tn = out["theta_need"].astype(np.float64)
    assert ((tn >= 0.25) & (tn <= 1.0)).all()
    cells = r["tree"].cells.astype(np.int64)
    # theta_need is the acceptance boundary of the walk's opening test
    rng = np.random.default_rng(0)
    for k in rng.choice(len(tn), size=min(40, len(tn)), replace=False):
        g, c = out["gc"][k]
        pb, npg = r["groups"][g]
        gp = pos_ph[pb:pb + npg, :3]
        mn, mx = gp.min(0), gp.max(0)
        gc_, gs_ = np.float32(0.5) * (mx + mn), np.float32(0.5) * (mx - mn)
        dd = np.abs(gc_ - r["size"][c, :3]) - gs_
        dd = (dd + np.abs(dd)) * np.float32(0.5)
        ds = np.sqrt(float((dd.astype(np.float64) ** 2).sum()))
        l, s = float(r["ls"][c, 0]), float(r["ls"][c, 1])
        for th, accept in ((tn[k] * (1 + 1e-4), True), (tn[k] * (1 - 1e-4), False)):
            assert (ds >= l / th + s) == accept
        # pair error by an independent numpy FP64 evaluation (reference P2C form,
        # code/computeForces.cu:83-128: D0..D3)
        if k < 10:
            pj = r["pos_tree"][cells[c, 2]:cells[c, 3]].astype(np.float64)
            dr = pj[None, :, :3] - gp[:, None, :].astype(np.float64)
            r2 = (dr ** 2).sum(-1) + np.float64(np.float32(1e-4))
            a_dir = (pj[None, :, 3, None] * dr / r2[..., None] ** 1.5).sum(1)
            m = r["mono"][c].astype(np.float64)
            q = np.concatenate([r["q0"][c], r["q1"][c]]).astype(np.float64)
            Q = np.array([[q[0], q[3], q[4]], [q[3], q[1], q[5]], [q[4], q[5], q[2]]])
            x = gp.astype(np.float64) - m[:3]
            rr2 = (x ** 2).sum(1) + np.float64(np.float32(1e-4))
            ri = 1 / np.sqrt(rr2)
            D1, D2, D3 = -m[3] * ri ** 3, 3 * m[3] * ri ** 5, -15 * m[3] * ri ** 7
            qR = x @ Q
            qRR = (qR * x).sum(1)
            C = D1 + 0.5 * (D2 * np.trace(Q) + D3 * qRR)
            a_p2c = C[:, None] * x + D2[:, None] * qR
            e = (np.linalg.norm(a_p2c - a_dir, axis=1) / amag[pb:pb + npg]).max()
            assert abs(e - out["err"][k]) <= 1e-6 * e + 1e-12
            assert abs(out["target"][k] - (np.log2(tn[k]) - np.log2(max(e, 1e-12)) / 3)) < 1e-5

"""Training driver of the opening-angle network (no reference counterpart).

The network (8 -> 32 -> 32 -> 1, FP16 parameters, see include/bonsai2_amd.h)
is trained on pairs the method generates itself: for sampled groups of the
Plummer configurations N=2^16 (theta 0.4, eps 0.01) and N=2^22 (theta 0.5,
eps 0.005), every candidate cell whose acceptance boundary theta_need lies in
[theta_lo, theta_hi] is measured against the FP64 direct force of its
particles (bt_learn_collect), and the network regresses
log2(theta_need) - log2(e) / 3 on the GPU (bt_mlp_train).  A per-pair error
budget then sets the angle: theta = exp2(y + log2(budget) / 3).

    python -m bonsai2_amd.learn --out bonsai2_amd/theta_mlp.json

writes the parameters, the band and the calibrated budget; Treecode.use_learned_theta
reads them back.  The band used at inference may be wider than the band the
pairs were collected in (the network extrapolates smoothly past theta_hi);

    python -m bonsai2_amd.learn --recalibrate --theta-hi 1.2

keeps the trained parameters of --out and recalibrates the budget for the new
inference band.
"""
import argparse
import json
import os
import time

import numpy as np

from . import Treecode, plummer, halo_disk, clustered, ORDER_PH

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "theta_mlp.json")
SEED = 19810614

# (name, n, theta, eps, groups sampled (0 = all)) -- the two Plummer
# configurations the network is trained on
TRAIN_CONFIGS = [
    ("plummer_2p16", 1 << 16, 0.4, 0.01, 0),
    ("plummer_2p22", 1 << 22, 0.5, 0.005, 2048),
]


def make_ic(ic, n, seed):
    if ic == "plummer":
        return plummer(n, seed=seed)
    if ic == "clustered":
        return clustered(n, seed=seed)
    if ic == "halo_disk":
        return halo_disk(n, seed=seed)
    raise ValueError(ic)


def built_tree(pos, vel, theta, eps, nleaf=16, ncrit=64, level_split=5):
    t = Treecode(eps, theta, ncrit)
    t.alloc(len(pos))
    t.ptcl_h2d(pos, vel)
    t.buildTree(nleaf)
    t.computeMultipoles()
    t.makeGroups(level_split, ncrit)
    return t


def collect(n, theta, eps, ngroups_sample, seed, theta_lo, theta_hi, ic="plummer",
            max_pairs=1 << 23):
    """Training pairs of one configuration -> (features, targets, info)."""
    pos, vel = make_ic(ic, n, seed)
    t = built_tree(pos, vel, theta, eps)
    ng = t.counts()["ngroups"]
    rng = np.random.default_rng(seed + 7)
    if ngroups_sample and ngroups_sample < ng:
        gsel = np.sort(rng.choice(ng, size=ngroups_sample, replace=False)).astype(np.int32)
    else:
        gsel = np.arange(ng, dtype=np.int32)
    col = t.learn_collect(gsel, theta_lo, theta_hi, max_pairs)
    t.close()
    if col["n"] > max_pairs:
        raise RuntimeError("%d pairs exceed max_pairs = %d" % (col["n"], max_pairs))
    # the device appends pairs in a nondeterministic order: sort by (group, cell)
    order = np.lexsort((col["gc"][:, 1], col["gc"][:, 0]))
    info = dict(n=n, theta=theta, eps=eps, groups=int(len(gsel)), pairs=col["n"])
    return col["feat"][order], col["target"][order], info


def train(x, y, steps=64000, batch=8192, lr=3e-3, seed=1):
    """Trains on the GPU; returns (params, loss curve)."""
    t = Treecode(0.01, 0.5, 64)
    t.mlp_init(seed, float(np.mean(y)))
    curve = t.mlp_train(x, y, steps=steps, batch=batch, lr=lr, seed=seed)
    p = t.mlp_params()
    t.close()
    return p, curve


def walk_error(t, n, samples, seed):
    """p50/p99 relative force error of the current forces against FP64 direct sums."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=min(n, samples), replace=False)).astype(np.int32)
    d = t.direct_fp64(idx)
    a = t.acc(ORDER_PH)[idx, :3].astype(np.float64)
    e = np.linalg.norm(a - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
    return float(np.percentile(e, 50)), float(np.percentile(e, 99))


def evaluate(t, params, budget, theta_lo, theta_hi, n, samples=4096, seed=5):
    """(p50, p99, interaction flops) of the learned walk on a built tree."""
    t.set_mlp_params(params)
    t.set_opening(1, budget, theta_lo, theta_hi)
    st = t.computeForces()
    p50, p99 = walk_error(t, n, samples, seed)
    flops = (20.0 * st[0] + 64.0 * st[2]) * n
    t.set_opening(0)
    return p50, p99, flops


def calibrate(params, theta_lo, theta_hi, p99_target, n=1 << 22, theta=0.5, eps=0.005,
              seed=SEED + 1, samples=4096):
    """Largest budget (grid of 8 steps per decade) whose learned walk keeps
    p99 <= p99_target on a Plummer sphere of n particles."""
    pos, vel = plummer(n, seed=seed)
    t = built_tree(pos, vel, theta, eps)
    best, table = None, []
    for b in 10.0 ** np.arange(-6.5, -2.9, 0.125):
        p50, p99, fl = evaluate(t, params, b, theta_lo, theta_hi, n, samples)
        table.append(dict(budget=float(b), p50=p50, p99=p99, flops=fl))
        if p99 <= p99_target:
            best = float(b)
    t.close()
    return best, table


def save(path, params, budget, theta_lo, theta_hi, meta):
    with open(path, "w") as f:
        json.dump(dict(params=[float(v) for v in params], budget=budget, theta_lo=theta_lo,
                       theta_hi=theta_hi, meta=meta), f, indent=0)


def load(path=DEFAULT_PATH):
    """(params float32[1377], budget, theta_lo, theta_hi) of a trained network."""
    with open(path) as f:
        d = json.load(f)
    return (np.asarray(d["params"], np.float32), float(d["budget"]), float(d["theta_lo"]),
            float(d["theta_hi"]))


def recalibrate(path, theta_lo, theta_hi, p99_target):
    """Calibrates the budget of a trained network for another inference band."""
    with open(path) as f:
        d = json.load(f)
    params = np.asarray(d["params"], np.float32)
    meta = d["meta"]
    meta.setdefault("train_band", [d["theta_lo"], d["theta_hi"]])
    t0 = time.time()
    budget, table = calibrate(params, theta_lo, theta_hi, p99_target)
    for r in table:
        print("budget %.2e  p50 %.2e  p99 %.2e  flops %.3e" %
              (r["budget"], r["p50"], r["p99"], r["flops"]), flush=True)
    if budget is None:
        raise SystemExit("no budget reaches p99 <= %g" % p99_target)
    meta.update(calibration=table, p99_target=p99_target)
    save(path, params, budget, theta_lo, theta_hi, meta)
    print("band [%g, %g]: budget %.3e -> %s (%.1f s)" %
          (theta_lo, theta_hi, budget, path, time.time() - t0), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--steps", type=int, default=64000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--theta-lo", type=float, default=0.25)
    ap.add_argument("--theta-hi", type=float, default=1.0)
    ap.add_argument("--p99-target", type=float, default=7e-4)
    ap.add_argument("--recalibrate", action="store_true",
                    help="keep the parameters of --out; calibrate the budget for the band given")
    args = ap.parse_args()
    t0 = time.time()
    if args.recalibrate:
        recalibrate(args.out, args.theta_lo, args.theta_hi, args.p99_target)
        return
    xs, ys, infos = [], [], []
    for k, (name, n, theta, eps, ngs) in enumerate(TRAIN_CONFIGS):
        x, y, info = collect(n, theta, eps, ngs, SEED + k, args.theta_lo, args.theta_hi)
        info["name"] = name
        xs.append(x)
        ys.append(y)
        infos.append(info)
        print("collected %s: %d pairs from %d groups" % (name, info["pairs"], info["groups"]),
              flush=True)
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    params, curve = train(x, y, args.steps, args.batch, args.lr)
    print("trained on %d pairs: loss %.4f -> %.4f (target variance %.4f)" %
          (len(y), curve[0], float(np.mean(curve[-50:])), float(np.var(y))), flush=True)
    budget, table = calibrate(params, args.theta_lo, args.theta_hi, args.p99_target)
    for r in table:
        print("budget %.2e  p50 %.2e  p99 %.2e  flops %.3e" %
              (r["budget"], r["p50"], r["p99"], r["flops"]), flush=True)
    if budget is None:
        raise SystemExit("no budget reaches p99 <= %g" % args.p99_target)
    meta = dict(configs=infos, train_band=[args.theta_lo, args.theta_hi], steps=args.steps, batch=args.batch, lr=args.lr,
                loss_first=float(curve[0]), loss_last=float(np.mean(curve[-50:])),
                target_var=float(np.var(y)), calibration=table, p99_target=args.p99_target,
                seconds=time.time() - t0)
    save(args.out, params, budget, args.theta_lo, args.theta_hi, meta)
    print("budget %.3e -> %s (%.1f s)" % (budget, args.out, time.time() - t0), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: one full gravity step (buildTree -> computeMultipoles -> makeGroups
-> computeForces -> kick-drift-kick) per timed step on one MI355X.

Default workload: N = 2^24 Plummer halo (mass 0.8) + exponential disk
(mass 0.2, scale length 0.3, sech^2 height 0.03), float32, eps = 0.002,
NLEAF 16, NCRIT 64, levelSplit 5, dt = 1/128, the learned per-pair opening
angle (bonsai2_amd/theta_mlp.json) with every step re-checked against 256
FP64 direct sums and falling back to fixed theta = 0.4 if their p99 exceeds
1e-3.  Synthetic seeded input, resident on the device before timing starts.

Prints ONE JSON line.  value = force evaluations per second of the whole step
(N / step time).  Beside it, on the same final particle state: the fixed
theta = 0.4 walk and step (the comparison of the metric), and both walks'
p50/p99 relative force errors against FP64 direct sums.  The walk kernel's
roofline uses the reference's flop count (20 per P2P, 64 per P2C;
code/treecode.cu:150-158) over the walk's own duration, taken with HIP events
on the stream the kernels run on.  cpu_baseline = the OpenMP C++ oracle (a
transcription of the reference path, fixed theta) on the host cores: build,
multipoles and groups in full, the walk on a seeded sample of the groups,
scaled to all groups.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("force-evals/s at N=2^24, p99 rel force err ≤1e-3 (learned vs fixed θ); "
          "interactions/s")
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per walk launch from two rocprofv3 --pmc passes (FETCH_SIZE, then
# WRITE_SIZE) scaled by a 1 GiB device copy (FETCH_SIZE reads 0.5x, WRITE_SIZE
# 1.0x): profiles/hbm_traffic_halo_disk_2p24_pmc.md.  Counters cannot be read
# from inside the bench run, so only the measured default configurations carry a
# figure; every other configuration reports null.
WALK_TRAFFIC_BYTES = {("halo_disk_2p24", "learned"): 4.095e9,
                      ("halo_disk_2p24", "fixed"): 1.126e10}

WORKLOADS = {
    "halo_disk_2p24": dict(n=1 << 24, ic="halo_disk", theta=0.4, eps=0.002),
    "clustered_2p24": dict(n=1 << 24, ic="clustered", theta=0.5, eps=0.002),
    "ensemble_256": dict(n=0, ic="ensemble", theta=0.5, eps=0.01),
    "plummer_2p24": dict(n=1 << 24, ic="plummer", theta=0.5, eps=0.002),
    "plummer_2p22": dict(n=1 << 22, ic="plummer", theta=0.5, eps=0.005),
    "plummer_2p20": dict(n=1 << 20, ic="plummer", theta=0.5, eps=0.005),
}


def make_ic(bt, ic, n, seed):
    if ic == "plummer":
        return bt.plummer(n, seed=seed)
    if ic == "clustered":
        return bt.clustered(n, seed=seed)
    return bt.halo_disk(n, seed=seed)


def cpu_baseline(pos, nleaf, ncrit, theta, eps, level_split, sample_groups, seed=5):
    from oracle import oracle as O
    t0 = time.perf_counter()
    box, _, _ = O.bounding_box(pos)
    tree = O.build_tree(pos, nleaf, box)
    pos_tree = pos[tree.order]
    t1 = time.perf_counter()
    size, mono, q0, q1, _ = O.multipoles(tree.cells, pos_tree, theta)
    t2 = time.perf_counter()
    keys = O.ph_keys(pos_tree, box)
    perm = O.sort_keys(keys)
    pos_ph = pos_tree[perm]
    groups = O.make_groups(keys[perm], level_split, ncrit)
    t3 = time.perf_counter()
    rng = np.random.default_rng(seed)
    ns = min(len(groups), sample_groups)
    sel = np.sort(rng.choice(len(groups), size=ns, replace=False))
    O.walk(groups[sel], tree.level_range, pos_ph, pos_tree, tree.cells, size, mono, q0, q1,
           eps)
    t4 = time.perf_counter()
    frac = ns / len(groups)
    t_walk = (t4 - t3) / frac
    total = (t3 - t0) + t_walk
    return dict(value=len(pos) / total, unit="force-evals/s",
                cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                kind="port",
                sample="full build/multipoles/groups (%.1f s); walk on %d of %d groups "
                       "(seeded, %.2f%%, %.1f s) scaled x%.0f" % (
                           t3 - t0, ns, len(groups), 100 * frac, t4 - t3, 1 / frac),
                walk_s_scaled=t_walk, build_s=t1 - t0, multipoles_s=t2 - t1, groups_s=t3 - t2)


def bench_ensemble(args, bt):
    """256 independent Plummer systems, N_k = round(2^U[10, 16]), in one context:
    autotuned NLEAF / NCRIT / walk grid / stack, the step captured in a HIP
    graph and replayed; aggregate interactions/s and per-system energy drift."""
    seed = 19810614
    pos, vel, off = bt.ensemble(256, seed=seed)
    n = int(off[-1])
    w = WORKLOADS["ensemble_256"]
    t = bt.Treecode(w["eps"], w["theta"], 64)
    t.alloc(n, off)
    t.ptcl_h2d(pos, vel)
    tt0 = time.perf_counter()
    best, table = t.autotune(args.level_split)
    tune_s = time.perf_counter() - tt0
    nl, nc = best["nleaf"], best["ncrit"]
    t.buildTree(nl)
    t.computeMultipoles()
    t.makeGroups(args.level_split, nc)
    t.computeForces()
    ek0, ep0 = t.computeEnergies()
    steps = args.steps
    # eager steps for reference (and warm-up), then the captured step
    te0 = time.perf_counter()
    for _ in range(args.warmup):
        t.step(args.dt, nl, args.level_split, nc)
    ms_eager = (time.perf_counter() - te0) * 1e3 / max(args.warmup, 1)
    st, tot = t.walk_stats()
    t.step_graph(args.dt, nl, args.level_split, nc, nsteps=1)   # capture + first replay
    t0 = time.perf_counter()
    t.step_graph(args.dt, nl, args.level_split, nc, nsteps=steps)
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    st2, tot2 = t.walk_stats()
    ek1, ep1 = t.computeEnergies()
    e0 = ek0 + ep0
    e1 = ek1 + ep1
    drift = np.abs(e1 - e0) / np.abs(e0)
    inter = tot2[0] + tot2[1]
    out = {
        "metric": "interactions/s (ensemble of 256 systems, one persistent walk per step)",
        "value": inter / (ms_step * 1e-3),
        "unit": "interactions/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (256 seeded Plummer systems, resident in HBM)",
        "config": {"workload": "ensemble_256", "n": n, "nsys": 256, "theta": w["theta"],
                   "eps": w["eps"], "dt": args.dt, "level_split": args.level_split,
                   "autotuned": best, "hip_graph": True},
        "force_evals_per_s": n / (ms_step * 1e-3),
        "eager_ms_per_step": ms_eager,
        "autotune_s": tune_s,
        "autotune_table": table,
        "interactions_per_particle": {"p2p": st2[0], "p2c": st2[2]},
        "energy_drift": {"median": float(np.median(drift)), "p90": float(np.percentile(drift, 90)),
                         "max": float(drift.max()), "steps": args.warmup + steps + 1},
        "system_sizes": {"min": int(np.diff(off).min()), "max": int(np.diff(off).max()),
                         "total": n},
        "cpu_baseline": None,
    }
    t.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="halo_disk_2p24", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the particle count")
    ap.add_argument("--nleaf", type=int, default=16)
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--level-split", type=int, default=5)
    ap.add_argument("--dt", type=float, default=1.0 / 128)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-groups", type=int, default=32768)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kahan", action="store_true")
    ap.add_argument("--group-mode", type=int, default=0, choices=(0, 1, 2),
                    help="0 PH-key groups (makeGroups), 1 leaves, 2 tree cells <= ncrit")
    ap.add_argument("--walk-variant", type=int, default=0, choices=(0, 1),
                    help="0: j-data staged in LDS, 1: j-data by scalar loads")
    ap.add_argument("--opening", default="learned", choices=("learned", "fixed"))
    ap.add_argument("--mlp", default=None, help="trained network (default bonsai2_amd/theta_mlp.json)")
    ap.add_argument("--check-samples", type=int, default=256)
    ap.add_argument("--p99-max", type=float, default=1e-3)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device with torch (a plain copy kernel "
                         "of known size, for scaling HBM byte counters)")
    ap.add_argument("--monopole", action="store_true",
                    help="monopole-only P2C (the reference without QUADRUPOLE; 20 flops each)")
    ap.add_argument("--kahan-compare", action="store_true",
                    help="also report the fixed-theta walk with and without compensated sums")
    ap.add_argument("--compare-steps", type=int, default=3,
                    help="timed fixed-theta steps for the comparison (0 = walk only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        import torch
        src = torch.ones(args.calibrate_copy * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    import bonsai2_amd as bt
    if args.workload == "ensemble_256":
        bench_ensemble(args, bt)
        return
    w = dict(WORKLOADS[args.workload])
    if args.n:
        w["n"] = args.n
    n = w["n"]
    seed = 19810614
    pos, vel = make_ic(bt, w["ic"], n, seed)

    t = bt.Treecode(w["eps"], w["theta"], args.ncrit)
    if args.no_kahan:
        t.set_walk_options(kahan=0)
    t.set_walk_variant(args.walk_variant)
    t.set_group_mode(args.group_mode)
    t.alloc(n)
    t.ptcl_h2d(pos, vel)
    if args.monopole:
        args.opening = "fixed"
        t.set_quadrupole(False)
    flops_p2c = 20.0 if args.monopole else 64.0   # code/treecode.cu:149-153
    learned = args.opening == "learned"
    if learned:
        from bonsai2_amd import learn
        params, budget, th_lo, th_hi = learn.load(args.mlp or learn.DEFAULT_PATH)
        t.set_mlp_params(params)
        t.set_opening(1, budget, th_lo, th_hi)
        t.set_opening_check(args.check_samples, args.p99_max, seed)
    # first force evaluation (not timed), then warmup steps
    t.buildTree(args.nleaf)
    t.computeMultipoles()
    t.makeGroups(args.level_split, args.ncrit)
    t.computeForces()
    ek0, ep0 = t.computeEnergies()
    e_start = float(ek0.sum() + ep0.sum())
    for _ in range(args.warmup):
        t.step(args.dt, args.nleaf, args.level_split, args.ncrit)

    walk_ms, phase = [], []
    flops, inter = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.step(args.dt, args.nleaf, args.level_split, args.ncrit)  # synchronises at the end
        tm = t.timings()
        phase.append(tm)
        walk_ms.append(tm["walk"])
        _, tot = t.walk_stats()
        flops.append(20.0 * tot[0] + flops_p2c * tot[1])
        inter.append(tot[0] + tot[1])
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps

    walk_ms = np.array(walk_ms)
    flops = np.array(flops)
    inter = np.array(inter)
    achieved = float((flops / (walk_ms * 1e-3)).mean() / 1e12)
    chk = t.opening_check() if learned else None
    # accuracy against the FP64 direct sum on seeded samples (outside the timed region)
    rng = np.random.default_rng(seed + 1)
    idx = np.sort(rng.choice(n, size=min(n, args.samples), replace=False)).astype(np.int32)
    d = t.direct_fp64(idx)

    def errors():
        a = t.acc(bt.ORDER_PH)[idx, :3].astype(np.float64)
        e = np.linalg.norm(a - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
        return tuple(float(x) for x in np.percentile(e, [50, 99]))

    p50, p99 = errors()
    ek, ep = t.computeEnergies()
    e_end = float(ek.sum() + ep.sum())
    c = t.counts()
    st_main, _ = t.walk_stats()
    kahan_cmp = None
    if args.kahan_compare:
        # the fixed-theta walk on the final state, compensated and plain FP32 sums
        if learned:
            t.set_opening_check(0, args.p99_max, seed)
            t.set_opening(0)
        res = {}
        accs = {}
        for kahan in (1, 0):
            t.set_walk_options(kahan=kahan)
            t.computeForces()
            accs[kahan] = t.acc(bt.ORDER_PH)[:, :3].astype(np.float64)
            k50, k99 = errors()
            res["kahan" if kahan else "plain"] = {"p50_rel_err": k50, "p99_rel_err": k99,
                                                  "walk_ms": t.timings()["walk"]}
        res["rel_rms_difference"] = float(np.sqrt(((accs[1] - accs[0]) ** 2).sum() /
                                                  (accs[1] ** 2).sum()))
        t.set_walk_options(kahan=0 if args.no_kahan else 1)
        if learned:
            t.set_opening(1, budget, th_lo, th_hi)
            t.set_opening_check(args.check_samples, args.p99_max, seed)
        kahan_cmp = res
    compare = None
    if learned:
        # the metric's comparison on the same particles and tree: the fixed
        # theta = 0.4 walk (and, with --compare-steps, its full step)
        t.set_opening_check(0, args.p99_max, seed)
        lw = []
        for _ in range(3):
            t.computeForces()
            lw.append(t.timings()["walk"])
        t.set_opening(0)
        t.set_params(theta=w["theta"])
        t.computeMultipoles()
        fw = []
        for _ in range(3):
            st_f = t.computeForces()
            fw.append(t.timings()["walk"])
        f50, f99 = errors()
        compare = {"theta": w["theta"], "walk_ms": float(np.median(fw)),
                   "learned_walk_ms": float(np.median(lw)),
                   "walk_speedup": float(np.median(fw) / np.median(lw)),
                   "p50_rel_err": f50, "p99_rel_err": f99,
                   "interactions_per_particle": {"p2p": st_f[0], "p2c": st_f[2]},
                   "flops_per_launch": (20.0 * st_f[0] + 64.0 * st_f[2]) * n}
        # the fixed angle as accurate as the learned one: the largest theta on a
        # 0.025 grid whose p99 (same samples, same particles) is at or below the
        # learned walk's p99
        eq = {"theta": w["theta"], "walk_ms": compare["walk_ms"], "p99_rel_err": f99}
        for th in np.arange(w["theta"] + 0.025, 0.801, 0.025):
            t.set_params(theta=float(th))
            t.computeMultipoles()
            t.computeForces()
            _, e99 = errors()
            if e99 > p99:
                break
            ew = [t.timings()["walk"]]
            t.computeForces()
            ew.append(t.timings()["walk"])
            eq = {"theta": round(float(th), 3), "walk_ms": float(min(ew)), "p99_rel_err": e99}
        eq["walk_speedup"] = eq["walk_ms"] / float(np.median(lw))
        compare["equal_p99"] = eq
        t.set_params(theta=w["theta"])
        t.computeMultipoles()
        t.computeForces()
        if args.compare_steps > 0:
            tc0 = time.perf_counter()
            for _ in range(args.compare_steps):
                t.step(args.dt, args.nleaf, args.level_split, args.ncrit)
            ms_f = (time.perf_counter() - tc0) * 1e3 / args.compare_steps
            compare["ms_per_step"] = ms_f
            compare["value"] = n / (ms_f * 1e-3)
            compare["step_speedup"] = ms_f / ms_step

    phase_ms = {k: float(np.mean([p[k] for p in phase])) for k in phase[0]}
    out = {
        "metric": METRIC,
        "value": n / (ms_step * 1e-3),
        "unit": "force-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded %s, resident in HBM)" % w["ic"],
        "config": {"workload": args.workload, "n": n, "theta": w["theta"], "eps": w["eps"],
                   "nleaf": args.nleaf, "ncrit": args.ncrit, "level_split": args.level_split,
                   "dt": args.dt,
                   "opening": ("learned per-pair theta (budget %.3g, band [%g, %g]), FP64 check "
                               "of %d samples per step, fallback theta %g at p99 > %g" % (
                                   budget, th_lo, th_hi, args.check_samples, w["theta"],
                                   args.p99_max)) if learned else "fixed theta",
                   "kahan": not args.no_kahan, "quadrupole": not args.monopole,
                   "walk_jdata": "scalar" if args.walk_variant == 1 else "lds",
                   "groups": ("PH keys", "leaves", "tree cells")[args.group_mode]},
        "interactions_per_s": float((inter / (walk_ms * 1e-3)).mean()),
        "walk_ms": float(walk_ms.mean()),
        "force_evals_per_s_walk": float(n / (walk_ms.mean() * 1e-3)),
        "p50_rel_err": p50,
        "p99_rel_err": p99,
        "error_samples": int(len(idx)),
        "energy": {"kinetic": float(ek.sum()), "potential": float(ep.sum()),
                   "E_start": e_start, "E_end": e_end,
                   "rel_drift": abs(e_end - e_start) / abs(e_start),
                   "steps_integrated": args.warmup + args.steps},
        "tree": {"ncells": c["ncells"], "nlevels": c["nlevels"], "ngroups": c["ngroups"]},
        "phase_ms": phase_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": (WALK_TRAFFIC_BYTES.get((args.workload, args.opening))
                                 if not args.n and args.nleaf == 16 and args.ncrit == 64
                                 else None),
                     "kernel": "k_walk<learned>" if learned else "k_walk",
                     "flops_per_launch": float(flops.mean()),
                     "avg_launch_ms": float(walk_ms.mean())},
    }
    st = st_main
    out["interactions_per_particle"] = {"p2p": st[0], "p2c": st[2], "p2p_max": st[1],
                                        "p2c_max": st[3]}
    if learned:
        out["opening_check"] = chk
        out["fixed_theta"] = compare
    if kahan_cmp is not None:
        out["kahan_compare"] = kahan_cmp
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, args.nleaf, args.ncrit, w["theta"], w["eps"],
                                           args.level_split, args.cpu_groups)
    else:
        out["cpu_baseline"] = None
    t.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

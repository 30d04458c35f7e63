"""Benchmark: one full gravity step (buildTree -> computeMultipoles -> makeGroups
-> computeForces -> kick-drift-kick) per timed step on one MI355X.

Headline workload: N = 2^24 Plummer halo (mass 0.8) + exponential disk
(mass 0.2, scale length 0.3, sech^2 height 0.03), float32, eps = 0.002,
NLEAF 16, NCRIT 64, levelSplit 5, dt = 1/128, the learned per-pair opening
angle (bonsai2_amd/theta_mlp.json) with every step re-checked against 256
FP64 direct sums and falling back to fixed theta = 0.4 if their p99 exceeds
1e-3.  Synthetic seeded input, resident on the device before timing starts.

Prints ONE JSON line.  value = force evaluations per second of the whole step
(N / step time of the eager steps, each with its HIP-event phase times); the
same step replayed from its HIP graph is reported beside it.  Also in the
line, on the same final particle state: the fixed theta = 0.4 walk and step
(the comparison of the metric), both walks' p50/p99 relative force errors
against FP64 direct sums, the walk kernel's roofline (the reference's flop
count, 20 per P2P and 64 per P2C, code/treecode.cu:150-158, over the walk's
own HIP-event duration), the HBM fraction of the build, sort and multipole
phases (algorithmic bytes over their HIP-event times), the plain fixed-angle
walk on a Plummer sphere of 2^24, every other configuration of the
benchmark (C1-C4) with its errors, drift and timings, and the CPU baseline:
the OpenMP C++ oracle (a transcription of the reference path, fixed theta)
on the host cores, median of 3.
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
HBM_PEAK_GBS = 8000.0
SEED = 19810614

WORKLOADS = {
    "halo_disk_2p24": dict(n=1 << 24, ic="halo_disk", theta=0.4, eps=0.002),
    "clustered_2p24": dict(n=1 << 24, ic="clustered", theta=0.5, eps=0.002),
    "ensemble_256": dict(n=0, ic="ensemble", theta=0.5, eps=0.01),
    "plummer_2p24": dict(n=1 << 24, ic="plummer", theta=0.4, eps=0.002),
    "plummer_2p22": dict(n=1 << 22, ic="plummer", theta=0.5, eps=0.005),
    "plummer_2p16": dict(n=1 << 16, ic="plummer", theta=0.4, eps=0.01),
    "plummer_2p20": dict(n=1 << 20, ic="plummer", theta=0.5, eps=0.005),
}


def make_ic(bt, ic, n, seed):
    if ic == "plummer":
        return bt.plummer(n, seed=seed)
    if ic == "clustered":
        return bt.clustered(n, seed=seed)
    return bt.halo_disk(n, seed=seed)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# ---- host description and the CPU baseline -------------------------------------

def host_info():
    cpus = sorted(os.sched_getaffinity(0))
    model, cores = "unknown", set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in list(f) + [""]:
                if not line.strip():
                    if cur.get("processor") is not None and int(cur["processor"]) in cpus:
                        cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                        model = cur.get("model name", model)
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
    except OSError:
        pass
    return dict(model=model, logical_cpus=len(cpus), physical_cores=len(cores) or None,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                omp_proc_bind=os.environ.get("OMP_PROC_BIND"))


def oracle_pipeline_times(O, pos, nleaf, ncrit, theta, eps, level_split, sample_groups, seed):
    """One oracle pass: build, multipoles, groups in full, the walk on a
    seeded sample of groups (all when sample_groups is None) -> seconds."""
    t0 = time.perf_counter()
    box, _, _ = O.bounding_box(pos)
    tree = O.build_tree(pos, nleaf, box)
    pos_tree = pos[tree.order]
    size, mono, q0, q1, _ = O.multipoles(tree.cells, pos_tree, theta)
    keys = O.ph_keys(pos_tree, box)
    perm = O.sort_keys(keys)
    pos_ph = pos_tree[perm]
    groups = O.make_groups(keys[perm], level_split, ncrit)
    t1 = time.perf_counter()
    if sample_groups is None or sample_groups >= len(groups):
        sel = np.arange(len(groups))
    else:
        rng = np.random.default_rng(seed)
        sel = np.sort(rng.choice(len(groups), size=sample_groups, replace=False))
    O.walk(groups[sel], tree.level_range, pos_ph, pos_tree, tree.cells, size, mono, q0, q1, eps)
    t2 = time.perf_counter()
    return t1 - t0, t2 - t1, len(sel), len(groups)


def cpu_baseline(bt, head_pos, head_w, args):
    """The OpenMP oracle on the host, median of 3: C1 (Plummer 2^16) in full,
    and the headline particles with the walk on a seeded sample of groups,
    scaled to all groups."""
    from oracle import oracle as O
    O.lib()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    info = host_info()
    w1 = WORKLOADS["plummer_2p16"]
    p1, _ = make_ic(bt, "plummer", w1["n"], SEED)
    c1 = []
    for _ in range(3):
        b, wk, _, _ = oracle_pipeline_times(O, p1, 16, 64, w1["theta"], w1["eps"], 5, None, 5)
        c1.append(b + wk)
    head = []
    for k in range(3):
        b, wk, ns, ng = oracle_pipeline_times(O, head_pos, args.nleaf, args.ncrit, head_w["theta"],
                                              head_w["eps"], args.level_split, args.cpu_groups, 5)
        head.append((b + wk * ng / ns, b, wk, ns, ng))
    head.sort()
    tot, b, wk, ns, ng = head[1]
    n = len(head_pos)
    return dict(
        value=n / tot, unit="force-evals/s", cores=int(info["omp_num_threads"] or
                                                       info["logical_cpus"]),
        kind="port",
        sample="headline particles (N=%d, fixed theta %g): build, multipoles, groups in full "
               "(%.1f s), walk on %d of %d groups (seeded, %.1f%%, %.1f s) scaled x%.1f; "
               "median of 3" % (n, head_w["theta"], b, ns, ng, 100.0 * ns / ng, wk, ng / ns),
        host=info,
        runs_s=[h[0] for h in head],
        c1_full={"n": w1["n"], "theta": w1["theta"], "eps": w1["eps"],
                 "step_s_median": float(np.median(c1)), "runs_s": c1,
                 "force_evals_per_s": w1["n"] / float(np.median(c1))})


# ---- accuracy -----------------------------------------------------------------

def rel_errors(t, bt, idx):
    d = t.direct_fp64(idx)
    a = t.acc(bt.ORDER_PH)[idx, :3].astype(np.float64)
    e = np.linalg.norm(a - d[:, :3], axis=1) / np.linalg.norm(d[:, :3], axis=1)
    return float(np.percentile(e, 50)), float(np.percentile(e, 99))


def sample_idx(n, k, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(n, k), replace=False)).astype(np.int32)


def new_tree(bt, pos, vel, w, ncrit=64, nleaf=16, level_split=5, offsets=None):
    t = bt.Treecode(w["eps"], w["theta"], ncrit)
    t.alloc(len(pos), offsets)
    t.ptcl_h2d(pos, vel)
    t.buildTree(nleaf)
    t.computeMultipoles()
    t.makeGroups(level_split, ncrit)
    t.computeForces()
    return t


def walk_rate(t, reps=3):
    ws = []
    for _ in range(reps):
        st = t.computeForces()
        ws.append(t.timings()["walk"])
    _, tot = t.walk_stats()
    w = float(np.median(ws))
    fl = 20.0 * tot[0] + 64.0 * tot[1]
    return dict(walk_ms=w, flops_per_launch=fl, tflops=fl / (w * 1e-3) / 1e12,
                frac=fl / (w * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                interactions_per_particle={"p2p": st[0], "p2c": st[2]},
                interactions_per_s=(tot[0] + tot[1]) / (w * 1e-3))


# ---- the configurations -------------------------------------------------------------

def config_c1(bt):
    """C1: Plummer 2^16, theta 0.4, eps 0.01: one force evaluation and one
    leapfrog step at dt = 1/64, both against the FP64 direct sum over all
    particles."""
    w = WORKLOADS["plummer_2p16"]
    pos, vel = make_ic(bt, "plummer", w["n"], SEED)
    t = new_tree(bt, pos, vel, w)
    allidx = np.arange(w["n"], dtype=np.int32)
    r = {"n": w["n"], "theta": w["theta"], "eps": w["eps"]}
    r.update(walk_rate(t))
    p50, p99 = rel_errors(t, bt, allidx)
    ek0, ep0 = t.computeEnergies()
    t.step(1.0 / 64, 16, 5, 64)
    q50, q99 = rel_errors(t, bt, allidx)
    ek1, ep1 = t.computeEnergies()
    r.update(p50_rel_err=p50, p99_rel_err=p99, after_step={"p50_rel_err": q50, "p99_rel_err": q99,
             "dt": 1.0 / 64}, error_samples=w["n"], step_ms=sum(t.timings().values()),
             rel_energy_change=float(abs((ek1 + ep1).sum() - (ek0 + ep0).sum()) / abs((ek0 + ep0).sum())))
    # wall clock per step at this size (about 300 small launches per step):
    # eager steps, then the same step replayed from its HIP graph
    t.step(1.0 / 64, 16, 5, 64)
    t0 = time.perf_counter()
    for _ in range(10):
        t.step(1.0 / 64, 16, 5, 64)
    eager = (time.perf_counter() - t0) * 1e3 / 10
    t.step_graph(1.0 / 64, 16, 5, 64, nsteps=1)
    t0 = time.perf_counter()
    t.step_graph(1.0 / 64, 16, 5, 64, nsteps=10)
    graph = (time.perf_counter() - t0) * 1e3 / 10
    r["wall_ms_per_step"] = {"eager": eager, "graph": graph}
    t.close()
    return r


def config_c2(bt):
    """C2: Plummer 2^22, theta 0.5, eps 0.005: 64 leapfrog steps at dt = 1/128
    (captured step, replayed), |dE/E| in FP64, errors on 4096 samples."""
    w = WORKLOADS["plummer_2p22"]
    pos, vel = make_ic(bt, "plummer", w["n"], SEED + 2)
    t = new_tree(bt, pos, vel, w)
    r = {"n": w["n"], "theta": w["theta"], "eps": w["eps"]}
    r.update(walk_rate(t))
    ek0, ep0 = t.computeEnergies()
    t.step_graph(1.0 / 128, 16, 5, 64, nsteps=1)   # capture + first replay
    t0 = time.perf_counter()
    t.step_graph(1.0 / 128, 16, 5, 64, nsteps=63)
    ms = (time.perf_counter() - t0) * 1e3 / 63
    ek1, ep1 = t.computeEnergies()
    p50, p99 = rel_errors(t, bt, sample_idx(w["n"], 4096, SEED + 3))
    r.update(steps=64, dt=1.0 / 128, graph_ms_per_step=ms,
             rel_energy_drift=float(abs((ek1 + ep1).sum() - (ek0 + ep0).sum()) / abs((ek0 + ep0).sum())),
             p50_rel_err=p50, p99_rel_err=p99, error_samples=4096)
    t.close()
    return r


def config_c3(bt):
    """C3: clustered 2^24 (64 Plummer subclusters), theta 0.5, eps 0.002:
    compensated against plain FP32 accumulation, errors on 8192 samples."""
    w = WORKLOADS["clustered_2p24"]
    pos, vel = make_ic(bt, "clustered", w["n"], SEED + 4)
    t = new_tree(bt, pos, vel, w)
    idx = sample_idx(w["n"], 8192, SEED + 5)
    r = {"n": w["n"], "theta": w["theta"], "eps": w["eps"], "error_samples": 8192}
    accs = {}
    for kahan in (1, 0):
        t.set_walk_options(kahan=kahan)
        rate = walk_rate(t)
        p50, p99 = rel_errors(t, bt, idx)
        accs[kahan] = t.acc(bt.ORDER_PH)[:, :3].astype(np.float64)
        rate.update(p50_rel_err=p50, p99_rel_err=p99)
        r["kahan" if kahan else "plain"] = rate
    r["rel_rms_difference"] = float(np.sqrt(((accs[1] - accs[0]) ** 2).sum() /
                                            (accs[1] ** 2).sum()))
    r["nlevels"] = t.counts()["nlevels"]
    t.set_walk_options(kahan=1)
    t.close()
    return r


def config_c4(bt, args):
    """C4: 256 independent Plummer systems, N_k = round(2^U[10, 16]), theta
    0.5, eps 0.01: autotuned NLEAF / NCRIT / walk grid / stack, 32 steps of
    the captured step, aggregate interactions/s and per-system drift."""
    w = WORKLOADS["ensemble_256"]
    pos, vel, off = bt.ensemble(256, seed=SEED)
    n = int(off[-1])
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
    te0 = time.perf_counter()
    for _ in range(2):
        t.step(1.0 / 128, nl, args.level_split, nc)
    ms_eager = (time.perf_counter() - te0) * 1e3 / 2
    t.step_graph(1.0 / 128, nl, args.level_split, nc, nsteps=1)
    t0 = time.perf_counter()
    t.step_graph(1.0 / 128, nl, args.level_split, nc, nsteps=29)
    ms_step = (time.perf_counter() - t0) * 1e3 / 29
    st, tot = t.walk_stats()
    ek1, ep1 = t.computeEnergies()
    drift = np.abs((ek1 + ep1) - (ek0 + ep0)) / np.abs(ek0 + ep0)
    p50, p99 = rel_errors(t, bt, sample_idx(n, 4096, SEED + 6))
    r = {"n": n, "nsys": 256, "theta": w["theta"], "eps": w["eps"], "steps": 32,
         "autotuned": best, "autotune_s": tune_s, "graph_ms_per_step": ms_step,
         "eager_ms_per_step": ms_eager,
         "interactions_per_s": (tot[0] + tot[1]) / (ms_step * 1e-3),
         "force_evals_per_s": n / (ms_step * 1e-3),
         "interactions_per_particle": {"p2p": st[0], "p2c": st[2]},
         "energy_drift": {"median": float(np.median(drift)),
                          "p90": float(np.percentile(drift, 90)), "max": float(drift.max())},
         "p50_rel_err": p50, "p99_rel_err": p99, "error_samples": 4096,
         "system_sizes": {"min": int(np.diff(off).min()), "max": int(np.diff(off).max())}}
    t.close()
    return r


def plain_walk(bt):
    """The plain fixed-angle walk at N = 2^24 Plummer (theta 0.4, eps
    0.002): its HIP-event duration and the fraction of FP32 vector peak."""
    w = WORKLOADS["plummer_2p24"]
    pos, vel = make_ic(bt, "plummer", w["n"], SEED + 7)
    t = new_tree(bt, pos, vel, w)
    r = {"workload": "plummer_2p24", "n": w["n"], "theta": w["theta"], "eps": w["eps"],
         "kernel": "k_walk (fixed theta)"}
    r.update(walk_rate(t, reps=5))
    r["roofline"] = {"bound": "valu", "achieved": r["tflops"], "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": r["frac"]}
    p50, p99 = rel_errors(t, bt, sample_idx(w["n"], 4096, SEED + 8))
    r.update(p50_rel_err=p50, p99_rel_err=p99, error_samples=4096)
    t.close()
    return r


# ---- HBM fractions of the non-walk phases ------------------------------------------------

def hbm_phases(t, phase, n, sort_passes):
    """Algorithmic bytes of the build, sort and multipole phases over their
    HIP-event times (bytes per kernel as listed in DESIGN.md)."""
    tr = t.tree(multipoles=False)
    cells = tr["cells"].astype(np.int64)
    node = cells[:, 1] != 0xFFFFFFFF
    nsplit = int((cells[node, 3] - cells[node, 2]).sum())
    ncells = len(cells)
    ngroups = t.counts()["ngroups"]
    b = {"bbox_keys": 44.0 * n,
         "sort": 24.0 * n * sort_passes,
         "levels": 32.0 * nsplit + 16.0 * ncells,
         "multipoles": 16.0 * n + 112.0 * ncells,
         "ph_keys": 28.0 * n,
         "ph_sort": 24.0 * n * sort_passes,
         "groups": 36.0 * n + 8.0 * n + 8.0 * ngroups}
    out = {}
    for k, v in b.items():
        ms = phase.get(k, 0.0)
        gbs = v / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"bytes": v, "ms": ms, "GB_s": gbs,
                  "frac": gbs / HBM_PEAK_GBS if gbs else None}
    out["sort_passes"] = sort_passes
    return out


# ---- the headline ------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="halo_disk_2p24",
                    choices=[k for k in sorted(WORKLOADS) if k != "ensemble_256"])
    ap.add_argument("--n", type=int, default=0, help="override the particle count")
    ap.add_argument("--nleaf", type=int, default=16)
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--level-split", type=int, default=5)
    ap.add_argument("--dt", type=float, default=1.0 / 128)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-groups", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip C1-C4 and the plain Plummer walk")
    ap.add_argument("--no-kahan", action="store_true")
    ap.add_argument("--group-mode", type=int, default=0, choices=(0, 1, 2),
                    help="0 PH-key groups (makeGroups), 1 leaves, 2 tree cells <= ncrit")
    ap.add_argument("--walk-variant", type=int, default=0, choices=(0, 1, 2),
                    help="0: one i-particle per lane, LDS j-data; 1: scalar-load j-data; "
                         "2: two i-particles per lane, half a batch per half-wave")
    ap.add_argument("--opening", default="learned", choices=("learned", "fixed"))
    ap.add_argument("--mlp", default=None, help="trained network (default bonsai2_amd/theta_mlp.json)")
    ap.add_argument("--check-samples", type=int, default=256)
    ap.add_argument("--p99-max", type=float, default=1e-3)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device with torch (a plain copy kernel "
                         "of known size, for scaling HBM byte counters)")
    ap.add_argument("--monopole", action="store_true",
                    help="monopole-only P2C (the reference without QUADRUPOLE; 20 flops each)")
    ap.add_argument("--phase-steps", type=int, default=3,
                    help="steps after the timed ones with the phases in sequence, for "
                         "per-phase times and HBM fractions")
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
    w = dict(WORKLOADS[args.workload])
    if args.n:
        w["n"] = args.n
    n = w["n"]
    pos, vel = make_ic(bt, w["ic"], n, SEED)
    log("headline %s n=%d" % (args.workload, n))

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
        t.set_opening_check(args.check_samples, args.p99_max, SEED)
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
    chk_a = t.opening_check() if learned else None
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
    chk_b = t.opening_check() if learned else None
    # after a fallback the statistics are the fixed-angle walk's (several times
    # the learned walk's interactions) while the walk's events time the learned
    # walk: such a step stays out of the rates
    fell = [False] * len(flops)
    if learned and chk_b["fallbacks"] > chk_a["fallbacks"]:
        med = float(np.median(flops))
        fell = [f > 1.5 * med for f in flops]
    # the same step replayed from its HIP graph (one host synchronisation)
    t.step_graph(args.dt, args.nleaf, args.level_split, args.ncrit, nsteps=1)
    chk_g = t.opening_check() if learned else None
    tg0 = time.perf_counter()
    t.step_graph(args.dt, args.nleaf, args.level_split, args.ncrit, nsteps=args.steps)
    ms_graph = (time.perf_counter() - tg0) * 1e3 / args.steps
    chk_c = t.opening_check() if learned else None
    log("headline: %.2f ms/step eager, %.2f ms/step graph" % (ms_step, ms_graph))
    # each phase alone (outside the timed region): the timed steps run the
    # groups beside the level build and multipoles, which shares the GPU
    # between the two branches and inflates both branches' event times
    t.set_overlap(False)
    phase_serial = []
    for _ in range(args.phase_steps):
        t.step(args.dt, args.nleaf, args.level_split, args.ncrit)
        phase_serial.append(t.timings())
    t.set_overlap(True)

    keep = ~np.array(fell, bool)
    if not keep.any():
        keep[:] = True
    walk_ms = np.array(walk_ms)[keep]
    flops = np.array(flops)[keep]
    inter = np.array(inter)[keep]
    achieved = float((flops / (walk_ms * 1e-3)).mean() / 1e12)
    chk = t.opening_check() if learned else None
    # accuracy against the FP64 direct sum on seeded samples (outside the timed region)
    t.computeForces()
    idx = sample_idx(n, args.samples, SEED + 1)
    p50, p99 = rel_errors(t, bt, idx)
    ek, ep = t.computeEnergies()
    e_end = float(ek.sum() + ep.sum())
    c = t.counts()
    st_main, _ = t.walk_stats()
    phase_ms = {k: float(np.mean([p[k] for p in phase])) for k in phase[0]}
    phase_ms_serial = ({k: float(np.mean([p[k] for p in phase_serial])) for k in phase_serial[0]}
                       if phase_serial else None)
    hbm = hbm_phases(t, phase_ms_serial or phase_ms, n, sort_passes=bt.SORT_PASSES)
    compare = None
    if learned:
        # the metric's comparison on the same particles and tree: the fixed
        # theta = 0.4 walk (and, with --compare-steps, its full step)
        t.set_opening_check(0, args.p99_max, SEED)
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
        f50, f99 = rel_errors(t, bt, idx)
        compare = {"theta": w["theta"], "walk_ms": float(np.median(fw)),
                   "learned_walk_ms": float(np.median(lw)),
                   "walk_speedup": float(np.median(fw) / np.median(lw)),
                   "p50_rel_err": f50, "p99_rel_err": f99,
                   "interactions_per_particle": {"p2p": st_f[0], "p2c": st_f[2]},
                   "flops_per_launch": (20.0 * st_f[0] + 64.0 * st_f[2]) * n}
        compare["roofline_frac"] = compare["flops_per_launch"] / (compare["walk_ms"] * 1e-3) / \
            1e12 / FP32_PEAK_TFLOPS
        # the fixed angle as accurate as the learned one: the largest theta on a
        # 0.025 grid whose p99 (same samples, same particles) is at or below the
        # learned walk's p99
        eq = {"theta": w["theta"], "walk_ms": compare["walk_ms"], "p99_rel_err": f99}
        for th in np.arange(w["theta"] + 0.025, 0.801, 0.025):
            t.set_params(theta=float(th))
            t.computeMultipoles()
            t.computeForces()
            _, e99 = rel_errors(t, bt, idx)
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
    t.close()
    log("headline done")

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
                   "walk_variant": args.walk_variant,
                   "groups": ("PH keys", "leaves", "tree cells")[args.group_mode]},
        "graph_ms_per_step": ms_graph,
        "value_graph": n / (ms_graph * 1e-3),
        "interactions_per_s": float((inter / (walk_ms * 1e-3)).mean()),
        "walk_ms": float(walk_ms.mean()),
        "force_evals_per_s_walk": float(n / (walk_ms.mean() * 1e-3)),
        "p50_rel_err": p50,
        "p99_rel_err": p99,
        "error_samples": int(len(idx)),
        "energy": {"kinetic": float(ek.sum()), "potential": float(ep.sum()),
                   "E_start": e_start, "E_end": e_end,
                   "rel_drift": abs(e_end - e_start) / abs(e_start),
                   "steps_integrated": args.warmup + 2 * args.steps + 1 + args.phase_steps},
        "tree": {"ncells": c["ncells"], "nlevels": c["nlevels"], "ngroups": c["ngroups"]},
        # timed steps: ph_keys, ph_sort and groups (second stream) overlap
        # levels and multipoles, each timed with the other branch beside it
        "phase_ms": phase_ms,
        # the same phases run in sequence (bt_set_overlap off), after the timed
        # steps; phase_hbm divides algorithmic bytes by these times
        "phase_ms_serial": phase_ms_serial,
        "phase_hbm": hbm,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     # HBM bytes per launch need two rocprofv3 --pmc passes, which
                     # cannot run inside this process: profiles/ holds them
                     "traffic": None,
                     "kernel": "k_walk<learned>" if learned else "k_walk",
                     "flops_per_launch": float(flops.mean()),
                     "avg_launch_ms": float(walk_ms.mean()),
                     "launches": int(keep.sum())},
    }
    st = st_main
    out["interactions_per_particle"] = {"p2p": st[0], "p2c": st[2], "p2p_max": st[1],
                                        "p2c_max": st[3]}
    if learned:
        # checks and fallbacks to the fixed angle inside the timed eager steps
        # and inside the graph replays (a fallback adds a fixed-angle walk)
        chk["timed_steps"] = {"checks": chk_b["checks"] - chk_a["checks"],
                              "fallbacks": chk_b["fallbacks"] - chk_a["fallbacks"]}
        chk["graph_steps"] = {"checks": chk_c["checks"] - chk_g["checks"],
                              "fallbacks": chk_c["fallbacks"] - chk_g["fallbacks"]}
        out["opening_check"] = chk
        out["fixed_theta"] = compare
    if not args.no_configs:
        configs = {"C5_halo_disk_2p24": {"see": "headline"}}
        for name, fn in (("C1_plummer_2p16", lambda: config_c1(bt)),
                         ("C2_plummer_2p22", lambda: config_c2(bt)),
                         ("C3_clustered_2p24", lambda: config_c3(bt)),
                         ("C4_ensemble_256", lambda: config_c4(bt, args))):
            log(name)
            configs[name] = fn()
        out["configs"] = configs
        log("plain walk")
        out["plain_walk"] = plain_walk(bt)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(bt, pos, w, args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Docking throughput on one MI355X.

Workload (the batched screen): 1,024 seeded ligands, heavy atoms uniform in
10-60, rotatable bonds Poisson(5) capped at 15, docked into the seeded
1,800-heavy-atom receptor with an 8 A pocket; 26 A box at 0.375 A (71^3
points); exhaustiveness 8; FP32 served search, one launch.  One step = the
whole screen: Monte Carlo search of 8,192 chains, merge, refinement and FP64
rescoring.  The receptor maps are built before the timed region; the ligand
descriptions are host inputs (their upload is inside it).

Beside it: the plain per-ligand FP64 path (one ligand per launch, FP64) on
a bounded sample, and the CPU oracle (smina's search restated in C++) on
a bounded sample of the same ligands, one single-threaded ligand per core.

--flex: the flexible-receptor batch instead: 256 ligands (20-50 heavy
atoms, 3-12 torsions), each docked with four flexible side chains (2-4 chi
torsions) cut from a 4,000-heavy-atom receptor, 24 A box, Vina scoring with
the hydrophobic weight x 1.1, mixed precision (FP32 evaluation, FP64 search
state).  --serve: the serving workload (see serve()).
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-smina_amd"))

HBM_PEAK_GBS = 8000.0
F64_STEP_DIV = 16  # plain FP64 sample: this fraction of the Monte Carlo steps
CPU_STEP_DIV = 4   # CPU baseline sample: this fraction of one chain's steps
TRAFFIC_FILE = "profiles/mc_kernel_screen1024_traffic.md"
T_START = time.perf_counter()


def size_sample(flats, k):
    """k ligand indices spread over the batch's cost range: the ligands at
    the (j + 0.5)/k quantiles of heavy atoms x (1 + torsions), so that a
    bounded sample stands for small, median and large ligands alike"""
    cost = np.array([f.nheavy * (1 + f.ntors_total) for f in flats], dtype=np.float64)
    order = np.argsort(cost, kind="stable")
    k = max(1, min(k, len(flats)))
    return [int(order[min(len(order) - 1, int((j + 0.5) / k * len(order)))]) for j in range(k)]


def measured_traffic(nligands, exhaustiveness):
    """HBM bytes per Monte Carlo launch measured by rocprofv3 --pmc on the
    1,024-ligand, exhaustiveness-8 screen (the workload is seeded, so the
    launch is the same one); None for any other workload"""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if nligands != 1024 or exhaustiveness != 8 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def screen_ligands(sg, n, seed=2024):
    """heavy atoms U[10, 60], torsions min(Poisson(5), 15), at most heavy-3
    (a tree of h atoms has at most h-3 bonds between non-terminal atoms)"""
    rng = np.random.default_rng(seed)
    heavy = rng.integers(10, 61, n)
    tors = np.minimum(np.minimum(rng.poisson(5.0, n), 15), heavy - 3)
    return [sg.Mol.ligand(i, int(h), int(t)) for i, (h, t) in enumerate(zip(heavy, tors))], heavy, tors


def serve(sg, args, log):
    """Serving: 8 resident receptors (2,000-6,000 heavy atoms, 30 A boxes at
    0.375 A, 12 maps each), Poisson arrivals of docking jobs (ligands of
    8-80 heavy atoms, 0-20 torsions, exhaustiveness from {4, 8, 16, 32}),
    continuously batched over concurrent launches; reports served ligands/s
    and p50/p99 job latency (submit to result)."""
    rng = np.random.default_rng(4242)
    recs = [sg.Mol.receptor(100 + k, int(rng.integers(2000, 6001)), 8.0) for k in range(8)]
    gd = sg.grid_dims((0.0, 0.0, 0.0), (30.0, 30.0, 30.0))
    types = list(range(2, 14))
    # each launch's search runs in 8 parts, and every part reports (a long
    # launch otherwise prints nothing until its jobs complete)
    srv = sg.Server(workers=args.workers, max_batch=64, progress_chunks=8, fast=args.fast,
                    progress=lambda f: log(f"a serving launch: Monte Carlo {100 * f:.0f}% of steps"))
    t = time.perf_counter()
    ri = [srv.add_receptor(r, gd, types) for r in recs]
    log(f"serving: {len(recs)} receptors resident ({time.perf_counter() - t:.1f}s to build maps)")
    heavy = rng.integers(8, 81, args.jobs)
    tors = np.minimum(np.minimum(rng.poisson(6.0, args.jobs), 20), np.maximum(heavy - 3, 0))
    exh = rng.choice([4, 8, 16, 32], args.jobs)
    target = rng.integers(0, len(recs), args.jobs)
    ligs = [sg.Mol.ligand(5000 + i, int(h), int(tt)) for i, (h, tt) in enumerate(zip(heavy, tors))]
    gaps = rng.exponential(1.0 / args.rate, args.jobs)
    t0 = time.perf_counter()
    res = [None] * args.jobs

    def waiter(i, job):
        res[i] = srv.wait(job)
        log(f"job {i} done: {int(heavy[i])} heavy atoms, exhaustiveness {int(exh[i])}, "
            f"latency {res[i]['latency_s']:.1f}s, batch {res[i]['batch_size']}")

    waiters = []
    for i in range(args.jobs):
        due = t0 + float(np.sum(gaps[:i + 1]))
        while time.perf_counter() < due:
            time.sleep(min(0.01, due - time.perf_counter()))
        th = threading.Thread(target=waiter, args=(i, srv.submit(ri[int(target[i])], ligs[i], int(exh[i]))))
        th.start()
        waiters.append(th)
        if (i + 1) % 8 == 0:
            log(f"submitted {i + 1}/{args.jobs}")
    for th in waiters:
        th.join()
    wall = time.perf_counter() - t0
    lat = np.array([r["latency_s"] for r in res])
    line = {
        "metric": "served docking jobs: ligands/s and p50/p99 job latency" +
                  (", --fast search (not smina's)" if args.fast else ""),
        "value": args.jobs / wall, "unit": "ligands/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "fp32", "data": "synthetic (seeded receptors and ligands)",
        "config": {"workload": "serving: 8 resident receptors (2,000-6,000 heavy atoms, 30 A boxes at 0.375 A, "
                               "12 FP32 maps each), Poisson arrivals of ligands with 8-80 heavy atoms, 0-20 "
                               "torsions, exhaustiveness from {4, 8, 16, 32}",
                   "jobs": args.jobs, "rate_per_s": args.rate, "workers": args.workers, "fast": bool(args.fast)},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(lat.max())},
        "queued_s_p50": float(np.median([r["queued_s"] for r in res])),
        "mean_batch": float(np.mean([r["batch_size"] for r in res])),
        "evals_per_s": float(sum(r["evals"] for r in res) / wall),
        "wall_s": wall,
    }
    print(json.dumps(line), flush=True)


def single(sg, args, log):
    """One ligand (38 heavy atoms, 10 torsions) at exhaustiveness 128 into
    the 1,800-atom receptor, 22.5 A box, docked in FP64, FP32 and mixed
    precision.  Reports each run's time, the top-pose score drift and
    heavy-atom RMSD against FP64, and the per-evaluation error of FP32
    against FP64 on the device (the FP64 evaluation is bit-exact with the
    oracle) over chain-start confs and confs near the pocket."""
    rec = sg.Mol.receptor(1, 1800, 8.0)
    lig = sg.Mol.ligand(2, 38, 10)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (22.5, 22.5, 22.5))
    exh = 128 if args.exhaustiveness == 8 else args.exhaustiveness
    runs = {}
    for name, prec in (("fp64", sg.FP64), ("fp32", sg.FP32), ("mixed", sg.MIXED)):
        d = sg.Docking(rec, gd, precision=prec, exhaustiveness=exh, seed=1, progress_chunks=8,
                       progress=lambda f, name=name: log(f"{name}: Monte Carlo {100 * f:.0f}% of steps"))
        d.prepare([lig])
        t = time.perf_counter()
        r = d.dock([lig], top_coords=True)[0]
        runs[name] = (r, time.perf_counter() - t)
        log(f"{name}: {runs[name][1]:.1f}s, top {r['e'][0]:.4f} kcal/mol")
        del d
    r64 = runs["fp64"][0]

    def drift(r):
        x, y = r["coords"], r64["coords"]
        return {"top_score_minus_fp64": r["e"][0] - r64["e"][0],
                "top_pose_rmsd_vs_fp64": float(np.sqrt(np.mean(np.sum((x - y) ** 2, axis=1))))}
    # per-evaluation error of the FP32 evaluation
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sg.vina_scoring())
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(rec.xyz, rec.type)
    types = sorted({int(t) for t in lig.type[:lig.nmovable] if t > 1})
    f = sg.Flat(lig)
    gids = {prec: ctx.build_grids(rid, gd, types, prec) for prec in (sg.FP64, sg.FP32)}
    bids = {prec: ctx.upload_models([f], prec) for prec in (sg.FP64, sg.FP32)}
    # the minima an FP64 search saves (32 chains, 200 steps each)
    cs, ss = zip(*[f.make_chain(7000 + k, gd) for k in range(32)])
    cnt, _, oc, _, _ = ctx.run_mc(bids[sg.FP64], gids[sg.FP64], sg.mc_params(), [0] * 32, np.array(cs), np.array(ss),
                                  [200] * 32, [(25 + f.natoms) // 3] * 32, coord_stride=3 * f.nheavy)
    confs = np.array([oc[k, j, :f.conf_size] for k in range(32) for j in range(cnt[k])])
    out = {}
    for prec in (sg.FP64, sg.FP32):
        out[prec] = ctx.eval_confs(bids[prec], [0] * len(confs), confs, f.change_size, f.natoms, (1000.0,) * 3,
                                   grid_id=gids[prec])
    e64, g64, x64 = out[sg.FP64]
    e32, g32, _ = out[sg.FP32]
    # the reference's FP32 rule: >= 99% within 1e-4 relative, every
    # evaluation not flagged at a discontinuity within 1e-3
    from accuracy import discontinuity_flags, fp32_rule
    flags = discontinuity_flags(f, gd, rec, x64, 0)
    within_rule, worst_unflagged, nflag, _ = fp32_rule(e32, g32, list(zip(e64, g64, x64)), flags)
    de = np.abs(e32 - e64)
    rel_e = de / np.maximum(1.0, np.abs(e64))
    rel_g = np.max(np.abs(g32 - g64), axis=1) / np.maximum(1.0, np.max(np.abs(g64), axis=1))
    small = np.abs(e64) < 10
    ctx.close()
    t32 = runs["fp32"][1]
    line = {
        "metric": "single ligand at exhaustiveness 128: FP32 vs FP64 per-evaluation error and top-pose drift",
        "value": 1.0 / t32, "unit": "ligands/s", "n_gpus": 1, "higher_is_better": True, "dtype": "fp32",
        "data": "synthetic (seeded receptor and ligand)",
        "config": {"workload": f"one ligand (38 heavy atoms, 10 torsions) into the 1,800-atom receptor, 22.5 A box, "
                               f"exhaustiveness {exh}, one launch per precision", "exhaustiveness": exh},
        "seconds": {k: v[1] for k, v in runs.items()},
        "top_score": {k: v[0]["e"][0] for k, v in runs.items()},
        "fp32": drift(runs["fp32"][0]), "mixed": drift(runs["mixed"][0]),
        "per_evaluation": {"confs": f"{len(confs)} minima saved by 32 FP64 chains of 200 steps",
                           "max_abs_dE_below_10_kcal": float(de[small].max()) if small.any() else None,
                           "median_abs_dE": float(np.median(de)),
                           "within_1e-4_rel": within_rule,
                           "flagged_at_discontinuity": nflag,
                           "max_rel_unflagged": worst_unflagged,
                           "rule_met": bool(within_rule >= 0.99 and worst_unflagged <= 1e-3),
                           "max_rel_grad": float(rel_g.max())},
    }
    print(json.dumps(line), flush=True)


def copy_calibration():
    """A plain device copy of known size (three copies of 1 GiB through
    torch), the reference that FETCH_SIZE / WRITE_SIZE counts are scaled by
    to give the Monte Carlo launch's HBM bytes."""
    import torch
    n = 256 * 1024 * 1024
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 3 * n * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--ligands", type=int, default=1024)
    ap.add_argument("--exhaustiveness", type=int, default=8)
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="wall-time budget of the whole run: timed steps stop early (and say so) to stay within it")
    ap.add_argument("--fp64-sample", type=int, default=8,
                    help="ligands docked on the plain FP64 path (sampled across the size range)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="ligands for the CPU baseline (0: one per core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true", help="the serving workload instead of the screen")
    ap.add_argument("--jobs", type=int, default=32, help="serving: jobs")
    ap.add_argument("--rate", type=float, default=1.0, help="serving: Poisson arrival rate, jobs/s")
    ap.add_argument("--workers", type=int, default=8, help="serving: concurrent launches")
    ap.add_argument("--single", action="store_true",
                    help="one ligand at exhaustiveness 128 in FP64, FP32 and mixed: error and top-pose drift")
    ap.add_argument("--flex", action="store_true",
                    help="the flexible-receptor batch (256 ligands, 4 side chains, mixed precision)")
    ap.add_argument("--minimize", type=int, default=0,
                    help="--minimize throughput on this many parsed poses (FP64, spline approximation)")
    ap.add_argument("--accuracy", type=int, default=0,
                    help="top-pose accuracy on this many ligands: served vs FP64, mixed vs FP64, FP64 seed 2 vs 1")
    ap.add_argument("--precision", choices=["fp32", "mixed", "fp64"], default=None,
                    help="search precision (default: fp32 for the screen, mixed for --flex)")
    ap.add_argument("--fast", action="store_true",
                    help="the screen with --fast's GPU-style search (automatic chain split and step fraction, early "
                         "stopping, two-pass screening), compared with the default search on the same ligands")
    ap.add_argument("--no-compare", action="store_true",
                    help="with --fast: skip the default-search run that the fast screen is compared with")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-plan autotuner (default plans)")
    ap.add_argument("--tune-steps", type=int, default=512,
                    help="autotuner: Monte Carlo steps of the longest chain per probe (the others scaled alike)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 3 x 1 GiB device copy (the reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    args.ligands_set = any(a.startswith("--ligands") for a in sys.argv[1:])
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the docking path runs on one "
              "GPU; ligand jobs would shard across GPUs with no collective on the critical path).", file=sys.stderr)
        sys.exit(2)

    if args.serve:
        # every worker context runs its size classes on four streams: give the
        # process a hardware queue per stream (set before the runtime starts;
        # streams that share a hardware queue run their kernels in FIFO order,
        # so a short job would wait behind another worker's long launch)
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, 4 * args.workers))
    import smina_gpu as sg

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.serve:
        serve(sg, args, log)
        return
    if args.single:
        single(sg, args, log)
        return
    if args.minimize > 0:
        minimize(sg, args, log)
        return

    if args.flex:
        rec0 = sg.Mol.receptor(7, 4000, 8.0)
        rng = np.random.default_rng(77)
        n = args.ligands if args.ligands_set else 256
        heavy = rng.integers(20, 51, n)
        tors = np.minimum(rng.integers(3, 13, n), heavy - 3)
        ligs, rigid = [], None
        for i, (h, t) in enumerate(zip(heavy, tors)):
            m, r = sg.Mol.flex(rec0, sg.Mol.ligand(9000 + i, int(h), int(t)), 7, 4, 2, 4)
            ligs.append(m)
            rigid = rigid or r
        rec, gd, hyd = rigid, sg.grid_dims((0.0, 0.0, 0.0), (24.0, 24.0, 24.0)), 1.1
        prec = {"fp32": sg.FP32, "mixed": sg.MIXED, "fp64": sg.FP64}[args.precision or "mixed"]
        workload = (f"flexible-receptor batch: {n} ligands (20-50 heavy atoms, 3-12 torsions) with 4 flexible side "
                    f"chains (2-4 chi torsions each) cut from a 4,000-atom receptor, 24 A box at 0.375 A, Vina "
                    f"scoring with hydrophobic x 1.1, exhaustiveness {args.exhaustiveness}, "
                    f"{args.precision or 'mixed'} precision, one launch")
        metric = "docked ligands/sec/GPU @ exh 8, flexible receptor"
    else:
        rec = sg.Mol.receptor(1, 1800, 8.0)
        ligs, heavy, tors = screen_ligands(sg, args.ligands)
        gd, hyd = sg.grid_dims((0.0, 0.0, 0.0), (26.0, 26.0, 26.0)), 1.0
        prec = {"fp32": sg.FP32, "mixed": sg.MIXED, "fp64": sg.FP64}[args.precision or "fp32"]
        workload = (f"batched screen: {len(ligs)} ligands (10-60 heavy atoms, Poisson(5) torsions <= 15) into a "
                    f"1,800-atom receptor, 26 A box at 0.375 A, exhaustiveness {args.exhaustiveness}, "
                    f"{args.precision or 'fp32'}, one launch")
        metric = "docked ligands/sec/GPU @ exh 8 (FP32 served vs FP64 plain); p50/p99 latency"
    flats = [sg.Flat(l) for l in ligs]
    log(f"inputs: {len(ligs)} models, {sum(f.npairs for f in flats)} pairs")

    # a long search reports its progress (smina's progress display): the
    # flexible batch runs its Monte Carlo steps as 8 consecutive launches
    chunks = 8 if args.flex else 1
    dock = sg.Docking(rec, gd, precision=prec, exhaustiveness=args.exhaustiveness, seed=1, hydrophobic_scale=hyd,
                      progress=lambda f: log(f"Monte Carlo {100 * f:.0f}% of steps") if chunks > 1 else None,
                      progress_chunks=chunks, fast=args.fast)
    if args.fast:
        workload += ", --fast"
        metric += ", --fast search (not smina's)"
    # build the maps for every ligand type before timing (receptor setup)
    dock.prepare(ligs)
    # warmup: small launches (a few ligands at a fraction of their steps)
    # that load the code objects and set up the device buffers; a full
    # screen as warmup would cost as much as a timed step
    warm = sg.Docking(rec, gd, precision=prec, exhaustiveness=args.exhaustiveness, seed=1, hydrophobic_scale=hyd,
                      step_fraction=1.0 / 64, fast=args.fast)
    warm_ligs = [ligs[i] for i in size_sample(flats, 4)]
    warm.prepare(warm_ligs)
    for w in range(args.warmup):
        warm.dock(warm_ligs)
    log(f"warmup: {args.warmup} small launch(es) of {len(warm_ligs)} ligands at 1/64 of their steps")
    del warm
    # the launch-plan autotuner on the screen's own chains (untimed): a few
    # Monte Carlo steps per candidate plan and size class; the timed steps
    # use the fastest plans
    tuned = None
    if not args.no_tune:
        t = time.perf_counter()
        tuned = dock.tune(ligs, probe_steps=args.tune_steps)
        log(f"autotuner: {time.perf_counter() - t:.1f}s, plans "
            f"{[(c['class'], c['chosen']['waves'], c['chosen']['persistent']) for c in tuned]}")
    # timed steps: each is the whole screen; run as many of the requested
    # steps as the time budget holds, leaving room for the baselines
    times = []
    results = None
    legs_s = 0.0 if args.no_cpu else 45.0
    legs_s += 60.0 if args.fp64_sample > 0 else 0.0
    for s in range(args.steps):
        elapsed = time.perf_counter() - T_START
        if times and elapsed + float(np.mean(times)) + legs_s > args.budget_s:
            log(f"time budget: {len(times)} of {args.steps} steps run ({elapsed:.0f}s elapsed, "
                f"{np.mean(times):.1f}s per step, budget {args.budget_s:.0f}s)")
            break
        t = time.perf_counter()
        results = dock.dock(ligs)
        times.append(time.perf_counter() - t)
        log(f"step {s + 1}/{args.steps}: {times[-1]:.2f}s")
    step_s = float(np.mean(times))
    value = len(ligs) / step_s
    mc_ms = dock.kernel_ms("mc")
    evals = np.array([r["evals"] for r in results], dtype=np.float64)
    nheavy = np.array([f.nheavy for f in flats], dtype=np.float64)
    npairs = np.array([f.npairs for f in flats], dtype=np.float64)
    # algorithmic bytes per evaluation: 8 corner values per heavy atom (one
    # corner-packed cell) + 2 table points of (e, dor) per pair, in the
    # evaluation's precision
    real = 8 if prec == sg.FP64 else 4
    bytes_per_eval = 8 * real * nheavy + 4 * real * npairs
    alg_bytes = float(np.sum(evals * bytes_per_eval))
    achieved = alg_bytes / (mc_ms / 1e3) / 1e9
    top = np.array([r["e"][0] if r["e"] else np.nan for r in results])
    traffic = None if args.flex or args.fast or prec != sg.FP32 else measured_traffic(len(ligs), args.exhaustiveness)
    fast = None
    if args.fast:
        fast = {} if args.no_compare else fast_vs_full(sg, args, log, rec, ligs, gd, hyd, prec, chunks, top)
        fast["full_searches"] = dock.full_searches()
    log(f"{'flex' if args.flex else 'screen'}: {value:.2f} ligands/s, MC kernel {mc_ms / 1e3:.2f}s, "
        f"{evals.sum():.3e} evals")

    # plain per-ligand FP64 path (one ligand per launch) on a bounded sample
    # across the batch's size range: 1/F64_STEP_DIV of the Monte Carlo steps,
    # the MC kernel time scaled back up (the search is step-linear)
    plain = None
    if args.fp64_sample > 0:
        frac = 1.0 / F64_STEP_DIV
        d64 = sg.Docking(rec, gd, fp64=True, exhaustiveness=args.exhaustiveness, seed=1, ligands_per_launch=1,
                         step_fraction=frac, hydrophobic_scale=hyd)
        pick = size_sample(flats, args.fp64_sample)
        sample = [ligs[i] for i in pick]
        d64.prepare(sample)
        per_lig = []
        for lig in sample:
            t = time.perf_counter()
            d64.dock([lig])
            dt = time.perf_counter() - t
            mc = d64.kernel_ms("mc") / 1e3
            per_lig.append(dt - mc + mc / frac)
        # the sample's mean time per ligand, weighted to the batch: each
        # sampled ligand stands for the ligands of its size stratum
        plain = {"value": len(sample) / float(np.sum(per_lig)), "unit": "ligands/s",
                 "sample": f"{len(sample)} ligands across the batch's size range (heavy atoms "
                           f"{[int(flats[i].nheavy) for i in pick]}, torsions {[int(flats[i].ntors_total) for i in pick]}), "
                           f"FP64, one ligand per launch, 1/{F64_STEP_DIV} of the MC steps with the MC kernel time "
                           f"scaled by {F64_STEP_DIV}",
                 "ligands": [int(i) for i in pick],
                 "seconds_per_ligand": per_lig}
        log(f"plain FP64: {plain['value']:.4f} ligands/s (extrapolated)")
        del d64

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sg, args, log, rec, ligs, flats, gd, hyd)

    accuracy = None
    if args.accuracy > 0:
        accuracy = top_pose_accuracy(sg, args, log, rec, ligs[:args.accuracy], gd, hyd, results[:args.accuracy])

    line = {
        "metric": metric,
        "value": value,
        "unit": "ligands/s",
        "n_gpus": 1,
        "steps": len(times),
        "steps_requested": args.steps,
        "warmup": args.warmup,
        "warmup_kind": "small launches: 4 ligands across the size range at 1/64 of their steps",
        "ms_per_step": step_s * 1e3,
        "step_ms": [t * 1e3 for t in times],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {sg.FP32: "fp32", sg.MIXED: "fp32 eval + fp64 search", sg.FP64: "fp64"}[prec],
        "data": "synthetic (seeded receptor and ligands)",
        "config": {"workload": workload, "ligands": len(ligs), "exhaustiveness": args.exhaustiveness,
                   "chains": len(ligs) * args.exhaustiveness,
                   "box_points": int((gd.n[0] + 1) * (gd.n[1] + 1) * (gd.n[2] + 1))},
        "evals_per_s": float(evals.sum() / (mc_ms / 1e3)),
        "latency_ms": {"p50": float(np.percentile(times, 50)) * 1e3, "p99": float(np.percentile(times, 99)) * 1e3,
                       "note": "per step over the steps run: a batched screen returns every ligand with its "
                               "launch, so a ligand's latency is its step's time; job latency under Poisson "
                               "arrivals is the serving run's (--serve)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "mc_kernel<float, float, false, false, true, false>" if prec == sg.FP32 else
                     "mc_kernel<float, double, false, false, true, false>" if prec == sg.MIXED else
                     "mc_kernel<double, double, false, false, true, false>",
                     "kernel_ms": mc_ms,
                     "algorithmic_bytes": alg_bytes,
                     "note": "one launch = the four size-class dispatches of one Monte Carlo launch, run "
                             "concurrently; kernel_ms spans them (HIP events on the launching stream); traffic "
                             "(calibrated FETCH_SIZE + WRITE_SIZE bytes per launch) from " + TRAFFIC_FILE +
                             " when this is the profiled workload, else null"},
        "cpu_baseline": cpu,
        "launch_plan": {"autotuned": tuned is not None, "classes": tuned,
                        "note": "per size class (pairs <= 256, <= 512, <= 1024, more): chains per workgroup and "
                                "one-shot or persistent work-queue grid, each candidate's probe time in ms"},
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / plain["value"]) if plain else None,
        "top_score_mean": float(np.nanmean(top)),
        "fp64_error_top_pose": fp64_error_summary(results),
    }
    if accuracy:
        line["top_pose_accuracy"] = accuracy
    if fast:
        # the Monte Carlo launch timed last is the second pass only, while
        # the evaluation counts span both passes: no roofline for this line
        line["roofline"] = None
        line["fast"] = fast
    print(json.dumps(line), flush=True)


def fp64_error_summary(results):
    """per-ligand FP64 error of an FP32 or mixed screen: each ligand's top
    pose, its refined energy in the search precision against the FP64 value
    of the same function at the same conf (None for an FP64 screen)"""
    err = np.array([r["fp64_error"] for r in results if r.get("fp64_error") is not None])
    if not len(err):
        return None
    rel = np.array([r["fp64_error"] / max(1.0, abs(r["e_fp64"])) for r in results
                    if r.get("fp64_error") is not None])
    return {"ligands": int(len(err)), "median_kcal": float(np.median(err)), "p99_kcal": float(np.percentile(err, 99)),
            "max_kcal": float(err.max()), "max_rel": float(rel.max()),
            "note": "|E_search - E_FP64| of each ligand's top pose: the refined energy in the search precision "
                    "against model::eval_deriv in FP64 (non_cache back end, bit-equal to the CPU oracle) at the "
                    "same conf"}


def fast_vs_full(sg, args, log, rec, ligs, gd, hyd, prec, chunks, top_fast):
    """the --fast screen's top scores against the default search's on the
    same ligands and seed (one untimed run): per-ligand differences, and how
    many of the default search's best 10% the fast screen also ranks there"""
    # a long default search reports its progress (8 parts)
    full = sg.Docking(rec, gd, precision=prec, exhaustiveness=args.exhaustiveness, seed=1, hydrophobic_scale=hyd,
                      progress=lambda f: log(f"default search: Monte Carlo {100 * f:.0f}% of steps"),
                      progress_chunks=8)
    full.prepare(ligs)
    t = time.perf_counter()
    out = full.dock(ligs)
    secs = time.perf_counter() - t
    top_full = np.array([r["e"][0] if r["e"] else np.nan for r in out])
    ok = np.isfinite(top_full) & np.isfinite(top_fast)
    d = top_fast[ok] - top_full[ok]
    q = max(1, int(np.ceil(0.1 * ok.sum())))
    best_full = set(np.argsort(np.where(ok, top_full, np.inf))[:q])
    best_fast = set(np.argsort(np.where(ok, top_fast, np.inf))[:q])
    log(f"default search: {len(ligs) / secs:.2f} ligands/s; fast minus default: median {np.median(d):.3f}, "
        f"top-10% overlap {len(best_full & best_fast)}/{q}")
    return {"default_ligands_per_s": len(ligs) / secs,
            "fast_minus_default": {"median": float(np.median(d)), "mean": float(np.mean(d)),
                                   "p90": float(np.percentile(d, 90)), "max": float(np.max(d)),
                                   "within_0.5": float(np.mean(d <= 0.5))},
            "top10_overlap": len(best_full & best_fast) / q,
            "note": "per-ligand top score (kcal/mol, after refinement and FP64 rescoring) of the --fast screen "
                    "minus the default search's, same seed; top10_overlap: share of the default search's best "
                    "10% that the fast screen also ranks in its best 10%"}


def minimize(sg, args, log):
    """--minimize (minimize_batch, src/gpu/gpu_docking.cpp:690-742): every
    pose refined in FP64 with the accurate line search and the force cap 10
    inside its own movable_atoms_box(4, 0.375), the spline approximation at
    factor 10 (the option's defaults), all poses in one launch.  The poses are the screen
    generator's ligands at their parsed coordinates (not docked: the
    reference's "random poses" row).  CPU baseline: the oracle's
    refine_structure on one core over a sample, each pose's result checked
    against the device's."""
    n = args.minimize
    rec = sg.Mol.receptor(1, 1800, 8.0)
    ligs, heavy, tors = screen_ligands(sg, n, seed=1890)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (26.0, 26.0, 26.0))
    d = sg.Docking(rec, gd, fp64=False, approximation="spline")
    d.minimize(ligs[:4], maxiters=0, accurate=True, force_cap=10.0)  # loads the code objects
    times, kms = [], []
    for k in range(max(1, args.steps)):
        t = time.perf_counter()
        res = d.minimize(ligs, maxiters=0, accurate=True, force_cap=10.0)
        times.append(time.perf_counter() - t)
        kms.append(d.kernel_ms("refine"))
        log(f"minimize {n} poses: {times[-1]:.3f}s (refine kernel {kms[-1]:.1f} ms)")
    es = np.array([r["e"] for r in res])
    wall = float(np.median(times))
    line = {"metric": "--minimize poses/s (FP64, spline approximation, one launch)", "value": n / wall,
            "unit": "poses/s", "n_gpus": 1, "steps": len(times), "warmup": 1, "ms_per_step": 1000 * wall,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded receptor and ligands at their parsed poses)",
            "config": {"workload": f"{n} poses (10-60 heavy atoms, Poisson(5) torsions <= 15) of the screen "
                                   f"generator, each in its movable_atoms_box(4, 0.375) against the 1,800-atom "
                                   f"receptor's direct terms, spline tables at factor 10, accurate line search, "
                                   f"force cap 10"},
            "refine_kernel_ms": float(np.median(kms)), "seconds": times,
            "energy_median": float(np.median(es[np.isfinite(es)])) if np.isfinite(es).any() else None}
    if not args.no_cpu:
        line["cpu_baseline"] = minimize_cpu_baseline(sg, args, log, rec, ligs, res)
    print(json.dumps(line), flush=True)


def minimize_cpu_baseline(sg, args, log, rec, ligs, res):
    """the oracle's refine_structure of a sample of the poses on one core,
    each in the box the device used; the tables are built once per pose box
    by the oracle and that set-up is left out, as smina's CPU --minimize
    builds them once per run"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    sf = sg.vina_scoring()
    pick = list(range(0, len(ligs), max(1, len(ligs) // 32)))[:32]
    spent, same = 0.0, 0
    for i in pick:
        r = res[i]
        g = sg.GridDims()
        for k in range(3):
            g.begin[k], g.end[k] = r["box"][0][k], r["box"][1][k]
            g.n[k] = int(round((g.end[k] - g.begin[k]) / 0.375))
        o = po.Oracle(sf, rec.xyz, rec.type, g)
        o.use_splines(10.0)
        m = o.add_model(ligs[i].desc_ptr)
        t = time.perf_counter()
        e, _, _ = o.refine(m, o.initial_conf(m), 10000, v=(10.0, 10.0, 10.0), accurate=True)
        spent += time.perf_counter() - t
        same += int(e == r["e"])
    log(f"CPU oracle --minimize: {len(pick)} poses in {spent:.2f}s on one core, {same} bit-equal to the device's")
    return {"value": len(pick) / spent, "unit": "poses/s", "cores": 1, "kind": "port",
            "sample": f"{len(pick)} poses spread over the set, one core, refine only (the oracle's per-box set-up "
                      f"left out); {same} of {len(pick)} FP64 energies bit-equal to the device's"}


def top_pose_accuracy(sg, args, log, rec, sample, gd, hyd, served):
    """per-ligand top score of the served search minus FP64's (the same
    seed, full search, one launch), beside FP64 at another seed minus FP64
    (the search's own noise floor), as the reference reports it"""
    def run(prec, seed):
        d = sg.Docking(rec, gd, precision=prec, exhaustiveness=args.exhaustiveness, seed=seed, hydrophobic_scale=hyd,
                       progress=lambda f: log(f"accuracy {prec_name[prec]} seed {seed}: {100 * f:.0f}%"),
                       progress_chunks=8)
        out = d.dock(sample)
        return np.array([r["e"][0] if r["e"] else np.nan for r in out])
    prec_name = {sg.FP32: "fp32", sg.MIXED: "mixed", sg.FP64: "fp64"}
    e64 = run(sg.FP64, 1)
    e64b = run(sg.FP64, 2)
    emix = run(sg.MIXED, 1)
    eserved = np.array([r["e"][0] if r["e"] else np.nan for r in served])

    def stats(d):
        d = d[np.isfinite(d)]
        return {"median": float(np.median(d)), "sd": float(np.std(d, ddof=1)), "max_abs": float(np.max(np.abs(d)))}
    acc = {"sample": f"first {len(sample)} ligands of the batch, exhaustiveness {args.exhaustiveness}, top score "
                     f"after refinement and FP64 rescoring, kcal/mol",
           "served_minus_fp64": stats(eserved - e64),
           "mixed_minus_fp64": stats(emix - e64),
           "fp64_seed2_minus_fp64": stats(e64b - e64)}
    log(f"top-pose accuracy: {json.dumps(acc)}")
    return acc


def cpu_baseline(sg, args, log, rec, ligs, flats, gd, hyd):
    """the FP64 oracle on the host cores: one ligand per core, its receptor
    maps plus one chain at 1/step_div of its Monte Carlo steps; the ligand's
    time is maps + exhaustiveness x step_div x the chain"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(cores, 16)
    nsample = min(args.cpu_sample or cores, len(ligs))
    pick = size_sample(flats, nsample)
    step_div = CPU_STEP_DIV * (16 if args.flex else 1)
    sf = sg.vina_scoring(hyd)
    params = sg.mc_params()

    def one(i):
        f = flats[i]
        o = po.Oracle(sf, rec.xyz, rec.type, gd)
        m = o.add_model(ligs[i].desc_ptr)
        t0 = time.perf_counter()
        o.build_grids(sorted(set(int(t) for t in ligs[i].type[:ligs[i].nmovable] if t > 1)))
        t1 = time.perf_counter()
        steps, maxit = sg.search_params(f.natoms, f.ntors_total)
        c0, s0 = f.make_chain(1000 + i, gd)
        _, _, _, ev = o.mc_chain(m, params, max(1, steps // step_div), maxit, c0, s0)
        t2 = time.perf_counter()
        return (t1 - t0) + args.exhaustiveness * step_div * (t2 - t1), ev, t2 - t1

    t = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        out = list(ex.map(one, pick))
    dt = time.perf_counter() - t
    est = [o[0] for o in out]
    cpu = {"value": cores / float(np.mean(est)), "unit": "ligands/s", "cores": cores, "kind": "port",
           "sample": f"{nsample} ligands across the batch's size range (heavy atoms "
                     f"{[int(flats[i].nheavy) for i in pick]}), one per core on {cores} threads: receptor maps plus "
                     f"one chain at 1/{step_div} of its Monte Carlo steps, per-ligand time = maps + "
                     f"{args.exhaustiveness} chains x {step_div} ({dt:.1f}s wall)",
           "evals_per_s_per_core": float(np.mean([o[1] / o[2] for o in out]))}
    log(f"CPU oracle: {cpu['value']:.3f} ligands/s on {cores} cores (extrapolated)")
    return cpu


if __name__ == "__main__":
    main()

"""Device parity against the CPU oracle (gfx950), through the C ABI.

FP64: bit-exact (maps, eval_deriv on both back ends, MC trajectories,
refinement, final rescoring, batched models).  FP32: the reference's rule
(src/gpu/tests/gpu_eval_test.cpp:55-92): at least 99% of evaluations within
1e-4 relative of the FP64 result for energy and gradient, and every
evaluation not flagged at a discontinuity (an atom on a map cell face or the
box edge, a receptor atom or an intramolecular pair at the 8 A cutoff,
flagged from the FP64 coordinates as :186-233 flags them) within 1e-3.  The
conformations are the reference's mix (:96-125): half fully random in the
box, half perturbed around the input pose.  Energies below 10 kcal/mol are
within 1e-3 kcal/mol of FP64 (the served tolerance)."""
import math

import numpy as np
import pytest

import smina_gpu as sg
from common import (discontinuity_flags, eval_confs_mix, fp32_rule, gd_c0, heavy_types, oracle_for, receptor,
                    random_confs)

pytestmark = pytest.mark.gpu

LIGANDS = [(1, 24, 6), (2, 38, 10), (4, 60, 15), (3, 10, 0)]


@pytest.fixture(scope="module")
def env():
    ligs = [sg.Mol.ligand(s, h, t) for s, h, t in LIGANDS]
    o, ids, sf, gd = oracle_for(ligs)
    rec = receptor()
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sf)
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(rec.xyz, rec.type)
    types = heavy_types(*ligs)
    gids = {p: ctx.build_grids(rid, gd, types, p) for p in (sg.FP64, sg.FP32)}
    flats = [sg.Flat(l) for l in ligs]
    bids = {p: ctx.upload_models(flats, p) for p in (sg.FP64, sg.FP32)}
    yield dict(ligs=ligs, o=o, ids=ids, sf=sf, gd=gd, ctx=ctx, rid=rid, types=types, gids=gids, flats=flats,
               bids=bids, rec=rec)
    ctx.close()


def test_device_present(env):
    name, cus = env["ctx"].device_info()
    print("device:", name, cus, "CUs", flush=True)
    assert "gfx950" in name and cus > 0


def test_grids_fp64_bit_exact(env):
    ctx, o = env["ctx"], env["o"]
    for t in env["types"]:
        assert np.array_equal(ctx.download_grid(env["gids"][sg.FP64], t, env["gd"]), o.grid(t))


def test_grids_offset_box_bit_exact(env):
    ctx = env["ctx"]
    gd = sg.grid_dims((0.1, 0.2, 0.3), (12.0, 12.0, 12.0))
    o, _, _, _ = oracle_for(env["ligs"][:1], gd=gd)
    types = heavy_types(env["ligs"][0])
    gid = ctx.build_grids(env["rid"], gd, types, sg.FP64)
    for t in types:
        assert np.array_equal(ctx.download_grid(gid, t, gd), o.grid(t))


def test_grids_fp32_are_rounded_fp64(env):
    ctx, o = env["ctx"], env["o"]
    for t in env["types"]:
        g = ctx.download_grid(env["gids"][sg.FP32], t, env["gd"])
        assert np.array_equal(g, o.grid(t).astype(np.float32).astype(np.float64))


def eval_both(env, li, prec, backend, v, n=24):
    ctx, o, f = env["ctx"], env["o"], env["flats"][li]
    confs = eval_confs_mix(f, env["gd"], n, seed=500 + 31 * li)
    kw = dict(grid_id=env["gids"][prec]) if backend == 0 else dict(grid_id=-1, receptor_id=env["rid"],
                                                                   gd=env["gd"], slope=10.0)
    e, g, x = ctx.eval_confs(env["bids"][prec], [li] * n, confs, f.change_size, f.natoms, v, **kw)
    ref = [o.eval_deriv(env["ids"][li], c, v, backend=backend, slope=1e6 if backend == 0 else 10.0)
           for c in confs]
    return confs, e, g, x, ref


@pytest.mark.parametrize("li", range(len(LIGANDS)))
@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("v", [10.0, 1000.0])
def test_eval_fp64_bit_exact(env, li, backend, v):
    print(f"eval fp64 ligand {LIGANDS[li]} backend {backend} v {v}", flush=True)
    confs, e, g, x, ref = eval_both(env, li, sg.FP64, backend, (v, v, v))
    for k, (eo, go, xo) in enumerate(ref):
        assert e[k] == eo, (k, e[k], eo)
        assert np.array_equal(g[k], go), (k, g[k] - go)
        assert np.array_equal(x[k], xo), k


@pytest.mark.parametrize("li", range(len(LIGANDS)))
@pytest.mark.parametrize("backend", [0, 1])
def test_eval_fp32_tolerance(env, li, backend):
    print(f"eval fp32 ligand {LIGANDS[li]} backend {backend}", flush=True)
    confs, e, g, x, ref = eval_both(env, li, sg.FP32, backend, (10.0, 10.0, 10.0), n=200)
    flags = discontinuity_flags(env["flats"][li], env["gd"], env["rec"], [r[2] for r in ref], backend)
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    print(f"  {within:.4f} within 1e-4, worst unflagged {worst:.2e} (conf {k}), {nflag} flagged", flush=True)
    assert within >= 0.99 and worst <= 1e-3, (within, worst, nflag, k, ref[k][0] if k >= 0 else None)
    for k, (eo, go, xo) in enumerate(ref):
        if abs(eo) < 10:
            assert abs(e[k] - eo) <= 1e-3, (k, e[k], eo)


def test_fp32_along_fp64_search(env):
    """The exhaustiveness-128 single-ligand check (38 heavy atoms, 10
    torsions): the poses an FP64 search saves (BFGS minima along 96 chains,
    about 900 poses) evaluated in FP32 and FP64 on the maps, at the hunt and
    authentic caps.  Reports |dE| and the gradient error.  The reference
    applies its rule (>= 99% within 1e-4 relative) to random confs
    (test_eval_fp32_tolerance holds it there); at minima the gradient is
    about 0, so the relative gradient error is the absolute one, where the
    FP32 coordinates alone (a few 1e-7 A) move it by about 1e-4: about 99%
    of minima are within 1e-4 (measured), and this test holds 98% and every
    unflagged pose within 1e-3, every unflagged energy below 10 kcal/mol
    within 1e-3 kcal/mol."""
    li = 1
    assert LIGANDS[li] == (2, 38, 10)
    ctx, f = env["ctx"], env["flats"][li]
    cs, ss, maxit, (cnt, e, c, x, ev) = run_chains(env, li, sg.FP64, 30, nchains=96, seed0=7000)
    confs = np.vstack([c[k, :cnt[k]] for k in range(len(cnt))])
    assert len(confs) >= 500
    n = len(confs)
    stats = []
    for v in (10.0, 1000.0):
        r = {}
        for prec in (sg.FP32, sg.FP64):
            r[prec] = ctx.eval_confs(env["bids"][prec], [li] * n, confs, f.change_size, f.natoms, (v, v, v),
                                     grid_id=env["gids"][prec])
        e32, g32, _ = r[sg.FP32]
        e64, g64, x64 = r[sg.FP64]
        flags = discontinuity_flags(f, env["gd"], env["rec"], x64, 0)
        _, worst, nflag, wk = fp32_rule(e32, g32, list(zip(e64, g64, x64)), flags)
        de = np.abs(e32 - e64)
        rel_e = de / np.maximum(1.0, np.abs(e64))
        rel_g = np.max(np.abs(g32 - g64), axis=1) / np.maximum(1.0, np.max(np.abs(g64), axis=1))
        within = np.mean(np.maximum(rel_e, rel_g) <= 1e-4)
        stats.append((v, n, float(de.max()), float(np.median(de)), float(rel_g.max()), float(within)))
        print(f"caps {v}: {n} poses, max |dE| {de.max():.2e}, median {np.median(de):.2e}, "
              f"max grad rel {rel_g.max():.2e}, within 1e-4: {within:.4f}, {nflag} flagged, worst unflagged "
              f"{worst:.2e}", flush=True)
        assert within >= 0.98 and worst <= 1e-3, (stats[-1], nflag, worst, wk)
        small = (np.abs(e64) < 10) & ~flags
        assert np.all(de[small] <= 1e-3), stats[-1]


def run_chains(env, li, prec, steps, nchains=4, seed0=2000):
    ctx, f = env["ctx"], env["flats"][li]
    cs, ss = zip(*[f.make_chain(seed0 + k, env["gd"]) for k in range(nchains)])
    maxit = (25 + f.natoms) // 3
    return cs, ss, maxit, ctx.run_mc(env["bids"][prec], env["gids"][prec], sg.mc_params(), [li] * nchains,
                                     np.array(cs), np.array(ss), [steps] * nchains, [maxit] * nchains,
                                     coord_stride=3 * f.nheavy)


@pytest.mark.parametrize("steps", [1, 10, 100])
def test_mc_fp64_trajectories_identical(env, steps):
    li = 0
    print(f"mc fp64 steps {steps}", flush=True)
    cs, ss, maxit, (cnt, e, c, x, ev) = run_chains(env, li, sg.FP64, steps)
    o, m = env["o"], env["ids"][li]
    for k in range(len(cs)):
        eo, co, xo, evo = o.mc_chain(m, sg.mc_params(), steps, maxit, cs[k], ss[k])
        assert cnt[k] == len(eo)
        assert np.array_equal(e[k, :cnt[k]], eo)
        assert np.array_equal(c[k, :cnt[k], :co.shape[1]], co)
        assert np.array_equal(x[k, :cnt[k], :xo.shape[1]], xo)
        assert ev[k] == evo


def test_mc_fp64_chunked_launches(env):
    """the search run as several launches (progress reporting): chains
    resume from their device state and still follow the oracle exactly"""
    ctx, o, li = env["ctx"], env["o"], 2
    seen = []
    ctx.set_progress(3, seen.append)
    try:
        cs, ss, maxit, (cnt, e, c, x, ev) = run_chains(env, li, sg.FP64, 25, nchains=3, seed0=2500)
    finally:
        ctx.set_progress(1)
    assert len(seen) == 3 and seen[-1] == 1.0 and seen == sorted(seen)
    for k in range(len(cs)):
        eo, co, xo, evo = o.mc_chain(env["ids"][li], sg.mc_params(), 25, maxit, cs[k], ss[k])
        assert cnt[k] == len(eo) and np.array_equal(e[k, :cnt[k]], eo)
        assert np.array_equal(c[k, :cnt[k], :co.shape[1]], co) and ev[k] == evo


def test_mc_fp64_early_stop(env):
    """early stopping (EarlyStop, stop_agreeing, run_monte_carlo;
    src/gpu/gpu_kernels.h:51-56, src/gpu/gpu_kernels_impl.cuh:607-657,
    :742-808): the chains run to first/n, (first+1)/n, .. of their steps; after
    each checkpoint but the last, a ligand whose units' best energies agree
    stops.  The expected stops come from the oracle's chains cut at each
    checkpoint (a chain's best energy is the lowest of its output
    container); every FP64 chain then matches the oracle's chain cut where
    its ligand stopped, poses and evaluation count alike."""
    ctx, o = env["ctx"], env["o"]
    n, first, fraction = 5, 2, 0.5
    chains = []  # (ligand, unit, conf, rng state, steps, maxiters)
    for li in (0, 2):
        f = env["flats"][li]
        for k in range(6):
            c, s = f.make_chain(7000 + 10 * li + k, env["gd"])
            chains.append((li, k // 2, c, s, 12 + 2 * li, (25 + f.natoms) // 3))
    ckpt = lambda steps, k: -(-steps * k // n)
    runs = {}  # (chain, checkpoint) -> oracle chain cut there
    for i, (li, u, c, s, steps, maxit) in enumerate(chains):
        for k in range(first, n + 1):
            runs[i, k] = o.mc_chain(env["ids"][li], sg.mc_params(), ckpt(steps, k), maxit, c, s)

    def expected_stop(eps):
        stop = {}
        for k in range(first, n):
            for li in (0, 2):
                if li in stop:
                    continue
                ub = {}
                for i, ch in enumerate(chains):
                    if ch[0] == li:
                        eo = runs[i, k][0]
                        ub[ch[1]] = min(ub.get(ch[1], np.inf), eo.min() if len(eo) else np.inf)
                best = min(ub.values())
                agree = sum(v - best <= eps for v in ub.values())
                if agree >= max(2, math.ceil(fraction * len(ub) - 1e-9)):
                    stop[li] = k
        return stop

    # the spread of the two best units at each checkpoint: an eps between
    # spreads makes a ligand stop at a later checkpoint than the first
    spreads = []
    for k in range(first, n):
        for li in (0, 2):
            ub = sorted(min(runs[i, k][0].min() for i, ch in enumerate(chains) if ch[0] == li and ch[1] == u)
                        for u in range(3))
            spreads.append(ub[1] - ub[0])
    cases = [0.0, 1e9] + sorted(set(spreads))[:2]
    confs = np.zeros((len(chains), 64))
    for i, ch in enumerate(chains):
        confs[i, :len(ch[2])] = ch[2]
    seen = set()
    for eps in cases:
        stop = expected_stop(eps)
        seen.add(tuple(sorted(stop.items())))
        print(f"early stop eps {eps:.6g}: stops {stop}", flush=True)
        p = sg.mc_params()
        p.early_stop_checkpoints, p.early_stop_first = n, first
        p.early_stop_fraction, p.early_stop_eps = fraction, eps
        cnt, e, c, x, ev = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], p, [ch[0] for ch in chains],
                                      confs, np.array([ch[3] for ch in chains]), [ch[4] for ch in chains],
                                      [ch[5] for ch in chains], units=[ch[1] for ch in chains])
        for i, ch in enumerate(chains):
            eo, co, xo, evo = runs[i, stop.get(ch[0], n)]
            assert cnt[i] == len(eo) and np.array_equal(e[i, :cnt[i]], eo), (eps, i)
            assert np.array_equal(c[i, :cnt[i], :co.shape[1]], co) and ev[i] == evo, (eps, i)
    assert len(seen) >= 2  # the cases stop differently


@pytest.mark.parametrize("li", [0, 1, 2])
def test_mc_fp64_visited_history(env, li):
    """QuickVina 2's history check (--gpu_qvina, src/gpu/gpu_mc.cuh:146-264):
    FP64 chains long enough for the ring of 5n points to fill and wrap
    follow the oracle's chains (QuickVina 2's own nearest-first form of the
    check) exactly, and skip local optimizations: fewer evaluations than
    the same chains without the history"""
    ctx, o, f = env["ctx"], env["o"], env["flats"][li]
    steps, n = 40, 4
    cs, ss = zip(*[f.make_chain(8100 + 10 * li + k, env["gd"]) for k in range(n)])
    maxit = (25 + f.natoms) // 3
    confs = np.zeros((n, 64))
    for k in range(n):
        confs[k, :len(cs[k])] = cs[k]
    p = sg.mc_params(visited=True)
    cnt, e, c, x, ev = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], p, [li] * n, confs, np.array(ss),
                                  [steps] * n, [maxit] * n, coord_stride=3 * f.nheavy)
    plain = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), [li] * n, confs, np.array(ss),
                       [steps] * n, [maxit] * n, coord_stride=3 * f.nheavy)[4]
    print(f"history: ligand {li}, n {f.change_size}, evals {list(ev)} vs {list(plain)} without", flush=True)
    for k in range(n):
        eo, co, xo, evo = o.mc_chain(env["ids"][li], p, steps, maxit, cs[k], ss[k])
        assert cnt[k] == len(eo) and np.array_equal(e[k, :cnt[k]], eo), k
        assert np.array_equal(c[k, :cnt[k], :co.shape[1]], co) and ev[k] == evo, k
    assert ev.sum() < plain.sum()


def test_mc_fp64_batched_models(env):
    """chains of different ligands in one launch follow their own CPU chains"""
    ctx, o = env["ctx"], env["o"]
    models, cs, ss, steps, iters = [], [], [], [], []
    for li, f in enumerate(env["flats"]):
        for k in range(2):
            c, s = f.make_chain(3000 + 10 * li + k, env["gd"])
            cc = np.zeros(64)
            cc[:f.conf_size] = c
            models.append(li)
            cs.append(cc)
            ss.append(s)
            steps.append(5 + li)
            iters.append((25 + f.natoms) // 3)
    cnt, e, c, x, ev = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), models, np.array(cs),
                                  np.array(ss), steps, iters)
    for i, li in enumerate(models):
        f = env["flats"][li]
        eo, co, xo, evo = o.mc_chain(env["ids"][li], sg.mc_params(), steps[i], iters[i], cs[i][:f.conf_size], ss[i])
        assert cnt[i] == len(eo) and np.array_equal(e[i, :cnt[i]], eo) and ev[i] == evo


def batched_chains(env, seed0=3000):
    models, cs, ss, steps, iters = [], [], [], [], []
    for li, f in enumerate(env["flats"]):
        for k in range(2):
            c, st = f.make_chain(seed0 + 10 * li + k, env["gd"])
            cc = np.zeros(64)
            cc[:f.conf_size] = c
            models.append(li)
            cs.append(cc)
            ss.append(st)
            steps.append(5 + li)
            iters.append((25 + f.natoms) // 3)
    return models, np.array(cs), np.array(ss), steps, iters


def check_against_oracle(env, models, cs, ss, steps, iters, out):
    cnt, e, c, x, ev = out
    for i, li in enumerate(models):
        f = env["flats"][li]
        eo, co, xo, evo = env["o"].mc_chain(env["ids"][li], sg.mc_params(), steps[i], iters[i], cs[i][:f.conf_size],
                                            ss[i])
        assert cnt[i] == len(eo) and np.array_equal(e[i, :cnt[i]], eo) and ev[i] == evo, (i, li)
        assert np.array_equal(c[i, :cnt[i], :co.shape[1]], co)


@pytest.mark.parametrize("form", ["graph-persistent", "graph-oneshot", "eager-persistent", "eager-oneshot",
                                  "graph-persistent-grid2", "eager-persistent-grid2",
                                  "graph-speculative4", "eager-speculative2", "graph-speculative8"])
def test_mc_fp64_launch_forms(env, form, monkeypatch):
    """the Monte Carlo launch in each of its forms: persistent grids over the
    device work queue (a grid of 2 workgroups makes each run several items),
    one workgroup per item, or one workgroup of 2 or 4 waves per chain
    (speculative line-search trials),
    the whole search captured in a HIP graph or enqueued directly, over
    three parts (progress reports as graph host nodes).  The graph is built
    and replayed on every call: two calls in a row with an eager call
    between them; every chain of every call follows the oracle exactly"""
    monkeypatch.setenv("SGPU_QUEUE", "1" if "persistent" in form else "0")
    monkeypatch.setenv("SGPU_SPEC", form[-1] if "speculative" in form else "1")
    if "grid2" in form:
        monkeypatch.setenv("SGPU_PERSIST_GRID", "2")
    if form.startswith("eager"):
        monkeypatch.setenv("SGPU_NO_GRAPH", "1")
    ctx = env["ctx"]
    models, cs, ss, steps, iters = batched_chains(env)
    seen = []
    ctx.set_progress(3, seen.append)
    try:
        for call in range(2):
            print(f"launch form {form}, call {call}", flush=True)
            out = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), models, cs, ss, steps, iters)
            check_against_oracle(env, models, cs, ss, steps, iters, out)
            with monkeypatch.context() as m:
                m.setenv("SGPU_NO_GRAPH", "1")
                eager = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), models, cs, ss, steps,
                                   iters)
            for u, w in zip(out, eager):
                assert np.array_equal(u, w)
    finally:
        ctx.set_progress(1)
    assert seen.count(1.0) == 4 and len(seen) == 12


def test_tune_mc_keeps_the_fastest_plan(env):
    """the launch-plan autotuner: joint candidates (the default plan, then
    a chains-per-workgroup target and one form, one-shot or persistent
    grid, for every size class at once) are timed on the whole request; the
    default is kept unless the fastest beats it by more than 3%; later
    searches with the tuned plans still follow the oracle exactly"""
    ctx = env["ctx"]
    models, cs, ss, steps, iters = batched_chains(env, seed0=3100)
    rep = ctx.tune_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), models, cs, ss, iters,
                      probe_steps=3)
    print("tuned:", rep, flush=True)
    assert len(rep) >= 2
    for cl in rep:
        ms = [cd["ms"] for cd in cl["candidates"]]
        assert len(ms) >= 2 and min(ms) > 0
        # one time per joint candidate, the same for every class
        assert ms == [cd["ms"] for cd in rep[0]["candidates"]]
        j = int(np.argmin(ms))
        best = cl["candidates"][j if ms[j] < ms[0] * (1 - 0.03) else 0]
        assert cl["chosen"] == {"waves": best["waves"], "persistent": best["persistent"]}
    out = ctx.run_mc(env["bids"][sg.FP64], env["gids"][sg.FP64], sg.mc_params(), models, cs, ss, steps, iters)
    check_against_oracle(env, models, cs, ss, steps, iters, out)


def test_fp32_search_is_deterministic(env):
    a = run_chains(env, 1, sg.FP32, 20)[3]
    b = run_chains(env, 1, sg.FP32, 20)[3]
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_refine_and_rescore_fp64_bit_exact(env):
    ctx, o = env["ctx"], env["o"]
    for li, f in enumerate(env["flats"]):
        confs = random_confs(f, env["gd"], 6, seed0=900 + li)
        confs[3:, :3] = 0.5
        maxit = (25 + f.natoms) // 3
        e, cr, xr = ctx.refine(env["bids"][sg.FP64], env["rid"], env["gd"], [li] * 6, confs, [maxit] * 6,
                               coord_stride=3 * f.nheavy)
        s0 = ctx.score(env["bids"][sg.FP64], env["rid"], env["gd"], [li] * 6, confs, 0)
        s1 = ctx.score(env["bids"][sg.FP32], env["rid"], env["gd"], [li] * 6, confs, 1)
        for k in range(6):
            eo, co, xo = o.refine(env["ids"][li], confs[k], maxit)
            assert e[k] == eo and np.array_equal(cr[k], co) and np.array_equal(xr[k], xo)
            assert s0[k] == o.eval(env["ids"][li], confs[k])
            assert s1[k] == o.eval_intramolecular(env["ids"][li], confs[k])


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("accurate", [False, True])
def test_run_bfgs_fp64_bit_exact(env, backend, accurate):
    """sgpu_run_bfgs (run_bfgs, src/gpu/gpu_kernels.h:104-111): one
    quasi_newton minimisation per conf on the maps or on the direct receptor
    term, for every ligand size; FP64 energies, confs and evaluation counts
    equal the oracle's bfgs (src/lib/bfgs.h:31-282) exactly; FP32 ends at a
    conf whose FP64 energy is within the served tolerance of its own"""
    ctx, o = env["ctx"], env["o"]
    models, confs, iters = [], [], []
    for li, f in enumerate(env["flats"]):
        c = eval_confs_mix(f, env["gd"], 4, seed=900 + li)
        for k in range(len(c)):
            row = np.zeros(max(x.conf_size for x in env["flats"]))
            row[:f.conf_size] = c[k]
            confs.append(row)
            models.append(li)
            iters.append((25 + f.natoms) // 3)
    confs = np.array(confs)
    v = (10.0, 10.0, 10.0) if backend == 0 else (1000.0, 1000.0, 1000.0)
    slope = 1e6 if backend == 0 else 10.0
    kw = dict(grid_id=env["gids"][sg.FP64]) if backend == 0 else dict(receptor_id=env["rid"], gd=env["gd"])
    e, c, ev = ctx.bfgs(env["bids"][sg.FP64], models, confs, iters, v, slope=slope, accurate=accurate, **kw)
    for k, li in enumerate(models):
        cs = env["flats"][li].conf_size
        eo, co, evo = o.bfgs(env["ids"][li], confs[k, :cs], v, iters[k], backend=backend, slope=slope,
                             accurate=accurate)
        assert e[k] == eo and np.array_equal(c[k, :cs], co) and ev[k] == evo, (k, li, e[k], eo, ev[k], evo)
    kw32 = dict(grid_id=env["gids"][sg.FP32]) if backend == 0 else kw
    e32, c32, _ = ctx.bfgs(env["bids"][sg.FP32], models, confs, iters, v, slope=slope, accurate=accurate, **kw32)
    for k, li in enumerate(models):
        cs = env["flats"][li].conf_size
        eo, _, xo = o.eval_deriv(env["ids"][li], c32[k, :cs], v, backend=backend, slope=slope)
        if discontinuity_flags(env["flats"][li], env["gd"], env["rec"], [xo], backend)[0]:
            continue
        assert abs(e32[k] - eo) <= 1e-3 * max(1.0, abs(eo) / 10), (k, li, e32[k], eo)


def test_atom_overlap_is_reported(env):
    """a ligand atom exactly on a receptor atom: the device reports it like
    the CPU (src/lib/non_cache.cpp:170-172), no result is returned"""
    ctx, f, rec = env["ctx"], env["flats"][0], env["rec"]
    heavy = np.where(rec.type > 1)[0]
    inside = heavy[np.all(np.abs(rec.xyz[heavy]) < 10, axis=1)][0]
    c = f.initial_conf()
    c[:3] = rec.xyz[inside]  # the root's first atom sits at the origin of the root frame
    with pytest.raises(sg.SgpuError) as err:
        ctx.eval_confs(env["bids"][sg.FP64], [0], c[None], f.change_size, f.natoms, (10, 10, 10), grid_id=-1,
                       receptor_id=env["rid"], gd=env["gd"], slope=10)
    assert err.value.status == 4


def test_invalid_requests_are_rejected(env):
    ctx, f = env["ctx"], env["flats"][0]
    with pytest.raises(sg.SgpuError):  # unknown batch
        ctx.eval_confs(12345, [0], f.initial_conf()[None], f.change_size, f.natoms, (10, 10, 10),
                       grid_id=env["gids"][sg.FP64])
    with pytest.raises(sg.SgpuError):  # model index out of range
        ctx.eval_confs(env["bids"][sg.FP64], [99], f.initial_conf()[None], f.change_size, f.natoms, (10, 10, 10),
                       grid_id=env["gids"][sg.FP64])
    with pytest.raises(sg.SgpuError):  # a map for a type the ligand lacks
        gid = ctx.build_grids(env["rid"], env["gd"], [2], sg.FP64)
        ctx.eval_confs(env["bids"][sg.FP64], [0], f.initial_conf()[None], f.change_size, f.natoms, (10, 10, 10),
                       grid_id=gid)


@pytest.mark.parametrize("mode", ["minimize-linear", "minimize-spline", "local_only"])
def test_minimize_fp64_matches_oracle(mode):
    """--minimize and --local_only (GpuDocking::minimize_batch): each parsed
    pose refined in FP64 inside its own movable_atoms_box, several boxes in
    one launch; bit-exact against the oracle's refine_structure over that
    box.  --minimize: the accurate line search, 10,000 iterations and a
    force cap of 10 (src/main/main.cpp:1216-1229), with the linear tables
    or the spline approximation, its default (factor 10; both sides
    evaluate precalculate_splines).  --local_only: the fast line search,
    setup_search's (25 + movable atoms) / 3 iterations and --force_cap's
    1000 (:513-522, :85)"""
    rec = receptor()
    ligs = [sg.Mol.ligand(40 + k, 16 + 6 * k, 2 + k, center=(0.5 * k, -0.5 * k, 0.25 * k)) for k in range(4)]
    approximation = "spline" if mode == "minimize-spline" else "linear"
    d = sg.Docking(rec, gd_c0(), fp64=False, approximation=approximation)
    local = mode == "local_only"
    cap = 1000.0 if local else 10.0
    if local:
        iters = [(25 + sg.Flat(lig).natoms) // 3 for lig in ligs]
        res = [d.minimize([lig], maxiters=it, accurate=False, force_cap=cap)[0] for lig, it in zip(ligs, iters)]
    else:
        iters = [10000] * len(ligs)
        res = d.minimize(ligs, maxiters=0, accurate=True, force_cap=cap)
    sf = sg.vina_scoring()
    import pyoracle as po
    for lig, r, it in zip(ligs, res, iters):
        gd = sg.GridDims()
        for i in range(3):
            gd.begin[i], gd.end[i] = r["box"][0][i], r["box"][1][i]
            gd.n[i] = int(round((gd.end[i] - gd.begin[i]) / 0.375))
        o = po.Oracle(sf, rec.xyz, rec.type, gd)
        if approximation == "spline":
            o.use_splines(10.0)
        m = o.add_model(lig.desc_ptr)
        e, _, _ = o.refine(m, o.initial_conf(m), it, v=(cap, cap, cap), accurate=not local)
        print(f"minimize ({mode}): e {r['e']:.6f} rmsd {r['rmsd']:.3f}", flush=True)
        assert r["e"] == e, (r["e"], e)


def test_spline_eval_bit_exact():
    """evaluation with the spline tables (precalculate_splines, factor 10):
    FP64 on the direct receptor term and on maps built from the splines
    equals the oracle's; FP32 on those maps stays within the served
    tolerance of it"""
    ligs = [sg.Mol.ligand(s, h, t) for s, h, t in LIGANDS[:3]]
    o, ids, sf, gd = oracle_for(ligs)
    o.use_splines(10.0)
    types = heavy_types(*ligs)
    o.build_grids(types)
    rec = receptor()
    ctx = sg.Context(0)
    try:
        nseg, coef, cut = sg.build_splines(sf, 10.0)
        ctx.upload_splines(nseg, coef, cut)
        rid = ctx.upload_receptor(rec.xyz, rec.type)
        flats = [sg.Flat(l) for l in ligs]
        bid = ctx.upload_models(flats, sg.FP64)
        bid32 = ctx.upload_models(flats, sg.FP32)
        g64 = ctx.build_grids(rid, gd, types, sg.FP64)
        g32 = ctx.build_grids(rid, gd, types, sg.FP32)
        v = (1000.0, 1000.0, 1000.0)
        direct = dict(grid_id=-1, receptor_id=rid, gd=gd, slope=10.0)
        for li, f in enumerate(flats):
            confs = random_confs(f, gd, 8, seed0=4100 + li)
            confs[4:, :3] = np.random.default_rng(li).uniform(-3, 3, (4, 3))
            e, g, x = ctx.eval_confs(bid, [li] * len(confs), confs, f.change_size, f.natoms, v, **direct)
            em, gm, xm = ctx.eval_confs(bid, [li] * len(confs), confs, f.change_size, f.natoms, v, grid_id=g64)
            e32, _, _ = ctx.eval_confs(bid32, [li] * len(confs), confs, f.change_size, f.natoms, v, grid_id=g32)
            for k in range(len(confs)):
                eo, go, _ = o.eval_deriv(ids[li], confs[k], v, backend=1, slope=10.0)
                assert e[k] == eo and np.array_equal(g[k], go), (li, k)
                eo, go, _ = o.eval_deriv(ids[li], confs[k], v, backend=0)
                assert em[k] == eo and np.array_equal(gm[k], go), (li, k)
                if abs(eo) < 1e3:
                    assert abs(e32[k] - eo) <= 1e-3, (li, k, e32[k], eo)
    finally:
        ctx.close()


def test_dock_fp64_splines_match_oracle():
    """a docking with --approximation spline (factor 32) in FP64: maps,
    search, refinement and rescoring all read the splines, and the modes
    equal the oracle's docking with the same splines"""
    rec, l = receptor(), sg.Mol.ligand(1, 24, 6)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=2, seed=5, step_fraction=0.05, approximation="spline",
                   approx_factor=32.0)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    o.use_splines(32.0)
    o.build_grids(heavy_types(l))
    e, c, ev = o.dock(ids[0], sg.mc_params(), seed=5, exhaustiveness=2, threads=2, step_fraction=0.05)
    print(f"dock with splines: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)


@pytest.mark.parametrize("lig,exh,frac,split,qv", [((3, 10, 0), 8, 1.0, 1, False), ((1, 24, 6), 4, 0.1, 1, False),
                                                   ((1, 24, 6), 2, 0.1, 4, False), ((1, 24, 6), 2, 0.1, 4, True)])
def test_dock_fp64_matches_oracle(lig, exh, frac, split, qv):
    """the single-ligand docking end to end through the host facade
    (GpuDocking::dock_batch: seeded chains, merge at 2 A, refinement,
    finish_docking's FP64 rescoring, mode selection) in FP64, against the
    oracle's dock (parallel_mc + refine_structure + finish_docking,
    src/lib/parallel_mc.cpp:80-98, src/main/main.cpp:256-324); the second
    case cuts every chain to a tenth of its steps on both sides, the third
    also splits every chain into 4 (--gpu_split: 4 x exhaustiveness chains
    with a quarter of those steps each, src/gpu/gpu_docking.cpp:200-220), the
    fourth adds QuickVina 2's history check (--gpu_qvina)"""
    rec, l = receptor(), sg.Mol.ligand(*lig)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=exh, seed=5, step_fraction=frac, chain_split=split,
                   skip_visited=qv)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    e, c, ev = o.dock(ids[0], sg.mc_params(visited=qv), seed=5, exhaustiveness=exh, threads=8, step_fraction=frac,
                      split=split)
    print(f"dock {lig} split {split} history {qv}: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)


def test_dock_fp64_force_cap_and_modes_match_oracle():
    """--force_cap and --num_modes through the host facade: the cap is
    main_procedure's authentic_v (src/main/main.cpp:392) of the refinement
    and the final scoring, while the chains keep monte_carlo's own 1000
    (src/lib/monte_carlo.cpp:96); setup_search keeps max(20, num_modes)
    minima per chain (:524).  FP64 at a tenth of the steps, against the
    oracle's dock with the same settings"""
    rec, l = receptor(), sg.Mol.ligand(1, 24, 6)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=7, step_fraction=0.1, num_modes=25,
                   energy_range=20.0, force_cap=50.0)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    e, c, ev = o.dock(ids[0], sg.mc_params(num_saved=25), seed=7, exhaustiveness=4, num_modes=25,
                      energy_range=20.0, threads=8, step_fraction=0.1, force_cap=50.0)
    print(f"dock with force cap 50 and 25 modes: {len(r['e'])} modes, {r['e'][:3]}", flush=True)
    assert len(r["e"]) > 9 and r["e"] == list(e)
    # the default cap differs: the setting reaches the refinement and scoring
    e1000 = o.dock(ids[0], sg.mc_params(num_saved=25), seed=7, exhaustiveness=4, num_modes=25, energy_range=20.0,
                   threads=8, step_fraction=0.1)[0]
    assert list(e1000) != list(e)


def test_dock_fp64_minimize_settings_match_oracle():
    """--minimize_iters and --minimize_early_term (every BFGS's iterations
    and early termination, the chains' and the refinement's,
    src/main/main.cpp:513-522, :1059-1066) through the host facade: FP64 at
    a tenth of the steps against the oracle's dock"""
    rec, l = receptor(), sg.Mol.ligand(3, 22, 5)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=13, step_fraction=0.1, minimize_iters=7,
                   minimize_early_term=True)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    p = sg.mc_params()
    p.early_term = 1
    e, c, ev = o.dock(ids[0], p, seed=13, exhaustiveness=4, threads=8, step_fraction=0.1, minimize_iters=7)
    print(f"dock with 7 iterations and early termination: {r['e'][:3]}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)
    e0 = o.dock(ids[0], sg.mc_params(), seed=13, exhaustiveness=4, threads=8, step_fraction=0.1)[0]
    assert list(e0) != list(e)


def test_dock_fp64_accurate_line_and_rmsd_filter_match_oracle():
    """--accurate_line (every BFGS, the chains' and the refinement's, with
    the accurate line search, src/main/main.cpp:1231-1234) and
    --min_rmsd_filter (remove_redundant's RMSD, :276) through the host
    facade: FP64 at a twentieth of the steps against the oracle's dock"""
    rec, l = receptor(), sg.Mol.ligand(2, 20, 4)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=11, step_fraction=0.05, num_modes=20,
                   energy_range=20.0, accurate_line=True, min_rmsd_filter=2.5)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    p = sg.mc_params()
    p.accurate_line_search = 1
    e, c, ev = o.dock(ids[0], p, seed=11, exhaustiveness=4, num_modes=20, energy_range=20.0, threads=8,
                      step_fraction=0.05, min_rmsd_filter=2.5)
    print(f"dock with the accurate line search, modes 2.5 A apart: {len(r['e'])} modes, {r['e'][:3]}", flush=True)
    assert len(r["e"]) > 1 and r["e"] == list(e)
    e1 = o.dock(ids[0], p, seed=11, exhaustiveness=4, num_modes=20, energy_range=20.0, threads=8,
                step_fraction=0.05)[0]
    assert len(e1) > len(e)  # the filter reaches the mode selection


@pytest.mark.parametrize("lig,exh", [((1, 24, 6), 4), ((3, 10, 0), 6)])
def test_dock_fp64_early_stop_matches_oracle(lig, exh):
    """--gpu_early_stop with --gpu_split 2 and a tenth of the steps through
    the host facade (3 checkpoints, the rule from the second; units of 2
    chains) in FP64, against the oracle's docking with the same rule"""
    rec, l = receptor(), sg.Mol.ligand(*lig)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=exh, seed=5, step_fraction=0.1, chain_split=2,
                   early_stop=True)
    r = d.dock([l])[0]
    o, ids, _, _ = oracle_for([l])
    p = sg.mc_params()
    p.early_stop_checkpoints, p.early_stop_first, p.early_stop_fraction, p.early_stop_eps = 3, 2, 0.5, 0.05
    e, c, ev = o.dock(ids[0], p, seed=5, exhaustiveness=exh, threads=8, step_fraction=0.1, split=2)
    full = o.dock(ids[0], sg.mc_params(), seed=5, exhaustiveness=exh, threads=8, step_fraction=0.1, split=2)[2]
    print(f"early stop {lig}: {r['e'][:3]}, oracle evals {ev} of {full} without stopping", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)


@pytest.mark.parametrize("prec", [sg.FP32, sg.MIXED], ids=["fp32", "mixed"])
def test_fp32_launch_forms_agree(env, prec, monkeypatch):
    """the served searches give the same chains in every launch form
    (persistent or one-shot grid, graph or direct launch): each kernel
    variant's copy of the evaluation rounds alike, so the launch plan the
    autotuner picks by timing never changes a result"""
    ctx = env["ctx"]
    bid = env["bids"][prec] if prec in env["bids"] else ctx.upload_models(env["flats"], prec)
    models, cs, ss, steps, iters = batched_chains(env, seed0=5300)
    monkeypatch.setenv("SGPU_SPEC", "1")
    outs = []
    for q, graph, grid in (("1", True, None), ("0", True, None), ("0", False, None), ("1", False, "2")):
        monkeypatch.setenv("SGPU_QUEUE", q)
        if grid:
            monkeypatch.setenv("SGPU_PERSIST_GRID", grid)
        else:
            monkeypatch.delenv("SGPU_PERSIST_GRID", raising=False)
        if graph:
            monkeypatch.delenv("SGPU_NO_GRAPH", raising=False)
        else:
            monkeypatch.setenv("SGPU_NO_GRAPH", "1")
        outs.append(ctx.run_mc(bid, env["gids"][sg.FP32], sg.mc_params(), models, cs, ss, steps, iters))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    if prec not in env["bids"]:
        ctx.release_models(bid)


@pytest.mark.parametrize("prec", [sg.FP32, sg.MIXED, sg.FP64], ids=["fp32", "mixed", "fp64"])
def test_speculative_chains_equal_sequential(env, prec, monkeypatch):
    """speculative line-search trials: with 2, 4 or 8 waves per chain (wave 0
    runs the chain, the others evaluate the next trial steps of its fast
    line search), every chain of a batch across the size classes gives the
    sequential search's result exactly: saved energies, confs, coordinates
    and evaluation counts, over both passes of each step (hunting and
    authentic caps).  In FP64 the chains also follow the oracle's exactly."""
    ctx = env["ctx"]
    bid = env["bids"][prec] if prec in env["bids"] else ctx.upload_models(env["flats"], prec)
    gid = env["gids"][sg.FP64 if prec == sg.FP64 else sg.FP32]
    models, cs, ss, steps, iters = batched_chains(env, seed0=5100)
    steps = [30 + 10 * li for li in models]
    outs = {}
    for w in ("1", "2", "4", "8"):
        monkeypatch.setenv("SGPU_SPEC", w)
        print(f"speculative chains: precision {prec}, width {w}", flush=True)
        outs[w] = ctx.run_mc(bid, gid, sg.mc_params(), models, cs, ss, steps, iters)
    for w in ("2", "4", "8"):
        for a, b in zip(outs["1"], outs[w]):
            assert np.array_equal(a, b), w
    assert min(outs["1"][0]) > 0
    if prec == sg.FP64:
        check_against_oracle(env, models, cs, ss, steps, iters, outs["8"])
    if prec not in env["bids"]:
        ctx.release_models(bid)


@pytest.mark.parametrize("prec", [sg.MIXED, sg.FP32])
def test_search_energies_match_fp64_oracle(env, prec):
    """the served searches: FP32 throughout, and mixed (FP32 evaluation,
    FP64 search state).  Every pose a Monte Carlo chain saves, and every
    refined pose, carries the energy the FP64 oracle gives its conf, within
    the served tolerance (1e-3 kcal/mol, 1e-4 relative above 10 kcal/mol)"""
    ctx, o = env["ctx"], env["o"]
    li = 1
    f = env["flats"][li]
    bid = ctx.upload_models(env["flats"], prec)
    cs, ss = zip(*[f.make_chain(6000 + k, env["gd"]) for k in range(4)])
    maxit = (25 + f.natoms) // 3
    cnt, e, c, x, ev = ctx.run_mc(bid, env["gids"][sg.FP32], sg.mc_params(), [li] * 4, np.array(cs), np.array(ss),
                                  [40] * 4, [maxit] * 4, coord_stride=3 * f.nheavy)
    worst = 0.0
    poses = []
    for k in range(4):
        assert cnt[k] > 0
        for j in range(cnt[k]):
            conf = c[k, j, :f.conf_size]
            eo = o.eval_deriv(env["ids"][li], conf, (1000.0,) * 3, backend=0, slope=1e6)[0]
            err = abs(e[k, j] - eo)
            assert err <= 1e-3 * max(1.0, abs(eo) / 10), (k, j, e[k, j], eo)
            worst = max(worst, err)
            poses.append(conf)
    confs = np.array(poses[:8])
    er, cr, xr = ctx.refine(bid, env["rid"], env["gd"], [li] * len(confs), confs, [maxit] * len(confs),
                            coord_stride=3 * f.nheavy)
    for k in range(len(confs)):
        if er[k] > 1e300:
            continue  # left the box at every slope
        # a refined pose lies inside the box, where the slope adds nothing
        eo = o.eval_deriv(env["ids"][li], cr[k], (1000.0,) * 3, backend=1, slope=10.0)[0]
        assert abs(er[k] - eo) <= 1e-3 * max(1.0, abs(eo) / 10), (k, er[k], eo)
    print(f"precision {prec}: worst saved-pose |dE| vs FP64 oracle {worst:.2e}", flush=True)
    ctx.release_models(bid)


@pytest.mark.parametrize("prec", [sg.FP32, sg.MIXED], ids=["fp32", "mixed"])
def test_dock_reports_fp64_error_per_result(prec):
    """every FP32 or mixed docking result carries the FP64 error of its top
    pose: the refined energy in the search precision (e_search) against the
    FP64 value of the refinement's own function at that conf (e_fp64, equal
    to the oracle's model::eval_deriv on the non_cache back end), within the
    served tolerance; an FP64 docking reports none"""
    rec = receptor()
    ligs = [sg.Mol.ligand(1, 24, 6), sg.Mol.ligand(3, 10, 0)]
    d = sg.Docking(rec, gd_c0(), precision=prec, exhaustiveness=2, seed=5, step_fraction=0.05)
    res = d.dock(ligs, confs=True)
    o, ids, _, _ = oracle_for(ligs)
    for k, r in enumerate(res):
        assert r["fp64_error"] is not None
        top = r["confs"][0]
        refs = [o.eval_deriv(ids[k], top, (1000.0,) * 3, backend=1, slope=10.0 ** p)[0] for p in range(1, 6)]
        print(f"ligand {k}: e_search {r['e_search']:.6f} e_fp64 {r['e_fp64']:.6f} error {r['fp64_error']:.2e}",
              flush=True)
        assert r["e_fp64"] in refs
        assert r["fp64_error"] == abs(r["e_search"] - r["e_fp64"])
        assert r["fp64_error"] <= 1e-3 * max(1.0, abs(r["e_fp64"]) / 10)
    d64 = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=1, seed=5, step_fraction=0.01)
    assert d64.dock(ligs[1:])[0]["fp64_error"] is None

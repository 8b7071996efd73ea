"""The screen (C2) and flexible-receptor (C3) configurations at their own
sizes, with the work cut to what the oracle finishes in seconds.

C2: the seeded 1,024-ligand set (heavy atoms U[10, 60], torsions
min(Poisson(5), 15), seed 2024) in the 26 A box at 0.375 A (71^3 points)
around the 1,800-atom receptor: 8 ligands across the set's size range share
one batch; FP64 maps and every FP64 chain (at 1/50 of its steps) equal the
oracle's, and the FP32 search's saved poses carry the oracle's FP64 energy
of their confs within the served tolerance.
C3: a 4,000-atom receptor with four flexible side chains (2-4 chi
torsions) cut near the pocket, the 24 A box (65^3 points), Vina with the
hydrophobic weight x 1.1: an FP64 docking at 1/20 of smina's steps equals
the oracle's docking exactly."""
import numpy as np
import pytest

import smina_gpu as sg
from common import discontinuity_flags, heavy_types, oracle_for

pytestmark = pytest.mark.gpu


def screen_set(n=1024, seed=2024):
    rng = np.random.default_rng(seed)
    heavy = rng.integers(10, 61, n)
    tors = np.minimum(np.minimum(rng.poisson(5.0, n), 15), heavy - 3)
    return heavy, tors


@pytest.fixture(scope="module")
def c2():
    heavy, tors = screen_set()
    cost = heavy * (1 + tors)
    order = np.argsort(cost, kind="stable")
    pick = [int(order[int((j + 0.5) / 8 * len(order))]) for j in range(8)]
    ligs = [sg.Mol.ligand(i, int(heavy[i]), int(tors[i])) for i in pick]
    rec = sg.Mol.receptor(1, 1800, 8.0)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (26.0, 26.0, 26.0))
    assert [gd.n[i] + 1 for i in range(3)] == [71, 71, 71]
    o, ids, sf, _ = oracle_for(ligs, gd=gd, rec=rec)
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sf)
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(rec.xyz, rec.type)
    types = heavy_types(*ligs)
    flats = [sg.Flat(l) for l in ligs]
    yield dict(ligs=ligs, pick=pick, rec=rec, gd=gd, o=o, ids=ids, ctx=ctx, rid=rid, types=types, flats=flats)
    ctx.close()


def c2_chains(S, seed0):
    models, cs, ss, steps, iters = [], [], [], [], []
    cw = max(f.conf_size for f in S["flats"])
    for li, f in enumerate(S["flats"]):
        full, maxit = sg.search_params(f.natoms, f.ntors_total)
        for k in range(2):
            c, st = f.make_chain(seed0 + 10 * li + k, S["gd"])
            cc = np.zeros(cw)
            cc[:f.conf_size] = c
            models.append(li)
            cs.append(cc)
            ss.append(st)
            steps.append(max(1, full // 50))
            iters.append(maxit)
    return models, np.array(cs), np.array(ss), steps, iters


def test_c2_maps_and_chains_fp64_bit_exact(c2):
    S = c2
    ctx, o = S["ctx"], S["o"]
    gid = ctx.build_grids(S["rid"], S["gd"], S["types"], sg.FP64)
    for t in S["types"]:
        assert np.array_equal(ctx.download_grid(gid, t, S["gd"]), o.grid(t)), t
    bid = ctx.upload_models(S["flats"], sg.FP64)
    models, cs, ss, steps, iters = c2_chains(S, 8000)
    print("C2 ligands", [(int(f.nheavy), int(f.ntors_total)) for f in S["flats"]], "steps", steps[::2], flush=True)
    cnt, e, c, x, ev = ctx.run_mc(bid, gid, sg.mc_params(), models, cs, ss, steps, iters,
                                  coord_stride=3 * max(f.nheavy for f in S["flats"]))
    for i, li in enumerate(models):
        f = S["flats"][li]
        eo, co, xo, evo = o.mc_chain(S["ids"][li], sg.mc_params(), steps[i], iters[i], cs[i][:f.conf_size], ss[i])
        assert cnt[i] == len(eo) and np.array_equal(e[i, :cnt[i]], eo) and ev[i] == evo, (i, li)


def test_c2_fp32_saved_poses_within_tolerance(c2):
    S = c2
    ctx, o = S["ctx"], S["o"]
    gid = ctx.build_grids(S["rid"], S["gd"], S["types"], sg.FP32)
    bid = ctx.upload_models(S["flats"], sg.FP32)
    models, cs, ss, steps, iters = c2_chains(S, 8100)
    cnt, e, c, x, ev = ctx.run_mc(bid, gid, sg.mc_params(), models, cs, ss, steps, iters,
                                  coord_stride=3 * max(f.nheavy for f in S["flats"]))
    worst, flagged, total = 0.0, 0, 0
    for i, li in enumerate(models):
        f = S["flats"][li]
        assert cnt[i] > 0
        for j in range(cnt[i]):
            eo, go, xo = o.eval_deriv(S["ids"][li], c[i, j, :f.conf_size], (1000.0,) * 3, backend=0, slope=1e6)
            err = abs(e[i, j] - eo)
            total += 1
            # the reference's rule leaves out evaluations at a discontinuity
            # (a map cell face, the cutoff, the box edge)
            if discontinuity_flags(f, S["gd"], S["rec"], [xo], 0)[0]:
                flagged += 1
                print(f"  flagged pose {i}/{j}: |dE| {err:.2e}", flush=True)
                continue
            worst = max(worst, err / max(1.0, abs(eo) / 10))
            assert err <= 1e-3 * max(1.0, abs(eo) / 10), (i, j, e[i, j], eo)
    print(f"C2 FP32 saved poses: {total}, {flagged} flagged, worst unflagged scaled |dE| {worst:.2e}", flush=True)


def test_c3_flex_dock_fp64_matches_oracle():
    rec4k = sg.Mol.receptor(7, 4000, 8.0)
    lig = sg.Mol.ligand(9000, 30, 6)
    model, rigid = sg.Mol.flex(rec4k, lig, 7, 4, 2, 4)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (24.0, 24.0, 24.0))
    assert [gd.n[i] + 1 for i in range(3)] == [65, 65, 65]
    o, ids, _, _ = oracle_for([model], gd=gd, rec=rigid, hydrophobic_scale=1.1)
    frac, exh = 0.05, 4
    d = sg.Docking(rigid, gd, fp64=True, exhaustiveness=exh, seed=3, hydrophobic_scale=1.1, step_fraction=frac)
    r = d.dock([model])[0]
    e, c, ev = o.dock(ids[0], sg.mc_params(), seed=3, exhaustiveness=exh, threads=exh, step_fraction=frac)
    print("C3 flex dock energies", r["e"], flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)

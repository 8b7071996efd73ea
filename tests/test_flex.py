"""The flexible-receptor configuration: a ligand plus four side chains with
2-4 chi torsions each, cut from the receptor (sgpu_synth_flex), scored with
the hydrophobic weight x 1.1.

CPU: the oracle's side-chain branches pinned by what the geometry fixes:
the zero conf reproduces the input, a chi torsion turns exactly the atoms
below it about its bond by the angle (right-handed, Rodrigues), the pair
list is exactly the heavy-atom pairs more than three bonds apart whose
distance can change, and gradients (side-chain torsions included) match
finite differences; the host's flattening and chain seeding equal the
oracle's.  GPU: FP64 evaluation, Monte Carlo chains, refinement, rescoring
and the whole docking pipeline bit-exact against the oracle; FP32 within
the served tolerance."""
import numpy as np
import pytest

import smina_gpu as sg
from common import (discontinuity_flags, eval_confs_mix, flex_model, fp32_rule, heavy_types, ligand, oracle_for,
                    random_confs, receptor)

V_DISTINCT = (10.0, 1000.0, 20.0)  # ligand pairs, receptor, other pairs


@pytest.fixture(scope="module")
def fx():
    model, rigid = flex_model()
    o, ids, sf, gd = oracle_for([model], rec=rigid, hydrophobic_scale=1.1)
    return dict(model=model, rigid=rigid, o=o, m=ids[0], sf=sf, gd=gd, flat=sg.Flat(model))


def desc_arrays(mol):
    d = mol.desc_ptr.contents
    nf, nt = d.nframes, d.ntrees
    get = lambda p, n: np.array([p[i] for i in range(n)])
    return dict(parent=get(d.frame_parent, nf), begin=get(d.frame_begin, nf), end=get(d.frame_end, nf),
                origin=get(d.frame_origin_atom, nf), axis=get(d.frame_axis_atom, nf),
                tree_frame=get(d.tree_frame, nt), bonds=get(d.bonds, 2 * d.nbonds).reshape(-1, 2),
                nligands=d.nligands, ntrees=nt)


def rodrigues(v, u, th):
    return v * np.cos(th) + np.cross(u, v) * np.sin(th) + np.outer(v @ u, u) * (1 - np.cos(th))


def coords(fx, c):
    return fx["o"].eval_deriv(fx["m"], c, (1000.0,) * 3, backend=1)[2].reshape(-1, 3)


def test_flex_model_shape(fx):
    model, rigid, flat = fx["model"], fx["rigid"], fx["flat"]
    d = desc_arrays(model)
    assert d["ntrees"] == 1 + 4 and d["nligands"] == 1
    chis = [(d["tree_frame"][t + 1] if t + 1 < d["ntrees"] else len(d["parent"])) - d["tree_frame"][t]
            for t in range(1, d["ntrees"])]
    assert all(2 <= c <= 4 for c in chis)
    assert flat.conf_size == 7 + 6 + sum(chis) and flat.change_size == 6 + 6 + sum(chis)
    assert flat.view.ninflex == model.natoms - model.nmovable >= 4
    # every atom after the ligand's came out of the receptor
    assert rigid.natoms == receptor().natoms - (model.natoms - ligand().natoms)
    # each side-chain root turns about the bond from its fixed anchor
    for t in range(1, d["ntrees"]):
        f0 = d["tree_frame"][t]
        assert d["parent"][f0] == -1 and d["axis"][f0] >= model.nmovable


def test_flex_flattening_matches_oracle(fx):
    o, m, flat, gd = fx["o"], fx["m"], fx["flat"], fx["gd"]
    info = o.info(m)
    a, b, lst = o.pairs(m)
    assert flat.npairs == info["npairs"] and flat.nother == info["nother"] > 0
    assert np.array_equal(a, flat.pair_a) and np.array_equal(b, flat.pair_b)
    lig_atoms = ligand().natoms  # the ligand's atoms come first
    # ligand-internal pairs in the ligand's list, everything touching a side chain in the other list
    inside = (a < lig_atoms) & (b < lig_atoms)
    assert np.array_equal(lst == 0, inside) and np.all(lst[~inside] == 2)
    assert flat.view.num_tors == int(o.num_tors(m))  # side chains do not count (src/lib/terms.cpp)
    assert np.array_equal(flat.initial_conf(), o.initial_conf(m))
    for seed in (1, 99, 123456):
        c1, s1 = flat.make_chain(seed, gd)
        c2, s2 = o.randomize(m, seed)
        assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
        # the side chains' torsions are drawn too
        assert np.all(c1[13:] != 0) and np.all(np.abs(c1[13:]) <= np.pi)


def test_flex_zero_conf_and_torsions_turn_the_side_chain(fx):
    model, flat = fx["model"], fx["flat"]
    d = desc_arrays(model)
    c0 = fx["o"].initial_conf(fx["m"])
    x0 = coords(fx, c0)
    assert np.max(np.abs(x0 - model.xyz[:model.nmovable])) < 1e-12
    frame_conf = np.ctypeslib.as_array(flat.view.frame_conf, (len(d["parent"]),))
    xyz_all = model.xyz
    rng = np.random.default_rng(5)
    checked = 0
    for t in range(1, d["ntrees"]):
        first = d["tree_frame"][t]
        last = d["tree_frame"][t + 1] if t + 1 < d["ntrees"] else len(d["parent"])
        for f in range(first, last):
            th = rng.uniform(-3, 3)
            c = c0.copy()
            c[frame_conf[f]] = th
            x = coords(fx, c)
            o_pos, a_pos = xyz_all[d["origin"][f]], xyz_all[d["axis"][f]]
            u = (o_pos - a_pos) / np.linalg.norm(o_pos - a_pos)
            moved = np.zeros(model.nmovable, bool)
            moved[d["begin"][f]:d["end"][last - 1]] = True  # this frame and the ones below it
            expect = x0.copy()
            expect[moved] = rodrigues(x0[moved] - o_pos, u, th) + o_pos
            assert np.max(np.abs(x - expect)) < 1e-12, (t, f)
            checked += 1
    assert checked >= 8


def test_flex_pairs_are_the_variable_distances(fx):
    """a pair is listed iff its atoms are heavy, more than three bonds
    apart, not both fixed, and their distance changes with the conf
    (src/lib/model.cpp:598-624 with the mobility of
    src/lib/parse_pdbqt.cpp:279-304).  One exception: the first atom of a
    side chain lies on its own rotation axis and never moves, but it belongs
    to a movable frame, so its pairs with atoms of other frames are listed
    although their distance is constant."""
    model, o, m, flat, gd = fx["model"], fx["o"], fx["m"], fx["flat"], fx["gd"]
    d = desc_arrays(model)
    n, nm = model.natoms, model.nmovable
    nb = [[] for _ in range(n)]
    for a, b in d["bonds"]:
        nb[a].append(b)
        nb[b].append(a)
    fixed = model.xyz[nm:]
    xs = [np.vstack([coords(fx, flat.make_chain(s, gd)[0]), fixed]) for s in (3, 4, 5)]
    listed = set(zip(flat.pair_a.tolist(), flat.pair_b.tolist()))
    still = np.all([np.all(x[:nm] == xs[0][:nm], axis=1) for x in xs], axis=0)
    frame = np.full(n, -1)
    for f in range(len(d["parent"])):
        frame[d["begin"][f]:d["end"][f]] = f
    assert np.array_equal(np.where(still)[0], np.sort(d["origin"][d["parent"] == -1][1:]))
    heavy = model.type > 1
    expect = set()
    for i in range(n):
        if not heavy[i]:
            continue
        dist = {i: 0}
        q = [i]
        while q:
            y = q.pop(0)
            if dist[y] == 3:
                continue
            for z in nb[y]:
                if z not in dist:
                    dist[z] = dist[y] + 1
                    q.append(z)
        for j in range(i + 1, n):
            if not heavy[j] or j in dist or (i >= nm and j >= nm):
                continue
            r = [np.linalg.norm(x[i] - x[j]) for x in xs]
            if max(r) - min(r) > 1e-9 or ((i < nm and still[i]) or (j < nm and still[j])) and frame[i] != frame[j]:
                expect.add((i, j))
    assert listed == expect


@pytest.mark.parametrize("backend", [0, 1])
def test_flex_gradient_matches_finite_differences(fx, backend):
    o, m, flat, gd = fx["o"], fx["m"], fx["flat"], fx["gd"]
    confs = random_confs(flat, gd, 4, seed0=81)
    h = 1e-5
    checked = flex_checked = 0
    for c in confs:
        c = c.copy()
        c[:3] = [0.5, -0.5, 1.0]
        e, g, x = o.eval_deriv(m, c, V_DISTINCT, backend=backend, slope=10)
        if not np.isfinite(e) or e > 1e4:
            continue
        x = x.reshape(-1, 3)
        if np.any(x < np.array(gd.begin) + 0.5) or np.any(x > np.array(gd.end) - 0.5):
            continue

        def fd_at(i, h):
            cp, cm = c.copy(), c.copy()
            if i < 3:
                cp[i] += h
                cm[i] -= h
            elif i < 6:
                from test_oracle_model import quat_mul, rot_quat
                v = np.zeros(3)
                v[i - 3] = h
                cp[3:7] = quat_mul(rot_quat(v), c[3:7])
                cm[3:7] = quat_mul(rot_quat(-v), c[3:7])
            else:
                cp[i + 1] += h
                cm[i + 1] -= h
            return (o.eval_deriv(m, cp, V_DISTINCT, backend=backend, slope=10)[0] -
                    o.eval_deriv(m, cm, V_DISTINCT, backend=backend, slope=10)[0]) / (2 * h)

        scale = max(1.0, np.max(np.abs(g)))
        for i in range(len(g)):
            a, b = fd_at(i, h), fd_at(i, h / 5)
            if abs(a - b) > 1e-3 * scale:
                continue  # a cutoff or a map cell face between the two steps
            assert abs(b - g[i]) / scale < 2e-2, (i, b, g[i])
            checked += 1
            flex_checked += i >= 12
    assert checked >= 20 and flex_checked >= 8, (checked, flex_checked)


# ------------------------------------------------------------------ device
@pytest.fixture(scope="module")
def dev(fx):
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(fx["sf"])
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(fx["rigid"].xyz, fx["rigid"].type)
    types = heavy_types(fx["model"])
    gids = {p: ctx.build_grids(rid, fx["gd"], types, p) for p in (sg.FP64, sg.FP32)}
    bids = {p: ctx.upload_models([fx["flat"]], p) for p in (sg.FP64, sg.FP32)}
    yield dict(ctx=ctx, rid=rid, gids=gids, bids=bids)
    ctx.close()


def eval_both(fx, dev, prec, backend, v, n):
    f, gd = fx["flat"], fx["gd"]
    confs = eval_confs_mix(f, gd, n, seed=700 + backend)
    kw = dict(grid_id=dev["gids"][prec]) if backend == 0 else dict(grid_id=-1, receptor_id=dev["rid"], gd=gd,
                                                                   slope=10.0)
    e, g, x = dev["ctx"].eval_confs(dev["bids"][prec], [0] * n, confs, f.change_size, f.natoms, v, **kw)
    ref = [fx["o"].eval_deriv(fx["m"], c, v, backend=backend, slope=1e6 if backend == 0 else 10.0) for c in confs]
    return e, g, x, ref


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("v", [V_DISTINCT, (1000.0, 1000.0, 1000.0)])
def test_flex_eval_fp64_bit_exact(fx, dev, backend, v):
    print(f"flex eval fp64 backend {backend} v {v}", flush=True)
    e, g, x, ref = eval_both(fx, dev, sg.FP64, backend, v, 24)
    for k, (eo, go, xo) in enumerate(ref):
        assert e[k] == eo, (k, e[k], eo)
        assert np.array_equal(g[k], go), (k, g[k] - go)
        assert np.array_equal(x[k], xo), k


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1])
def test_flex_eval_fp32_tolerance(fx, dev, backend):
    """the reference's FP32 rule (src/gpu/tests/gpu_eval_test.cpp:55-92):
    at least 99% within 1e-4 relative (energy and gradient), every
    evaluation not flagged at a discontinuity within 1e-3, on the
    reference's mix of random and perturbed confs; energies below 10
    kcal/mol within 1e-3 kcal/mol"""
    print(f"flex eval fp32 backend {backend}", flush=True)
    e, g, x, ref = eval_both(fx, dev, sg.FP32, backend, V_DISTINCT, 200)
    flags = discontinuity_flags(fx["flat"], fx["gd"], fx["rigid"], [r[2] for r in ref], backend)
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    print(f"  {within:.4f} within 1e-4, worst unflagged {worst:.2e} (conf {k}), {nflag} flagged", flush=True)
    assert within >= 0.99 and worst <= 1e-3, (within, worst, nflag, k)
    for k, (eo, go, xo) in enumerate(ref):
        if abs(eo) < 10:
            assert abs(e[k] - eo) <= 1e-3, (k, e[k], eo)


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 10, 60])
def test_flex_mc_fp64_trajectories_identical(fx, dev, steps):
    """Monte Carlo chains with side-chain torsion mutations
    (src/lib/mutate.cpp:25-63) follow the oracle's chains exactly"""
    print(f"flex mc fp64 steps {steps}", flush=True)
    f, o, m = fx["flat"], fx["o"], fx["m"]
    cs, ss = zip(*[f.make_chain(4000 + k, fx["gd"]) for k in range(4)])
    maxit = (25 + f.natoms) // 3
    cnt, e, c, x, ev = dev["ctx"].run_mc(dev["bids"][sg.FP64], dev["gids"][sg.FP64], sg.mc_params(), [0] * 4,
                                         np.array(cs), np.array(ss), [steps] * 4, [maxit] * 4,
                                         coord_stride=3 * f.nheavy)
    for k in range(4):
        eo, co, xo, evo = o.mc_chain(m, sg.mc_params(), steps, maxit, cs[k], ss[k])
        assert cnt[k] == len(eo)
        assert np.array_equal(e[k, :cnt[k]], eo)
        assert np.array_equal(c[k, :cnt[k], :co.shape[1]], co)
        assert np.array_equal(x[k, :cnt[k], :xo.shape[1]], xo)
        assert ev[k] == evo


@pytest.mark.gpu
def test_flex_refine_and_rescore_fp64_bit_exact(fx, dev):
    f, o, m, ctx = fx["flat"], fx["o"], fx["m"], dev["ctx"]
    confs = random_confs(f, fx["gd"], 6, seed0=950)
    confs[3:, :3] = 0.5
    maxit = (25 + f.natoms) // 3
    e, cr, xr = ctx.refine(dev["bids"][sg.FP64], dev["rid"], fx["gd"], [0] * 6, confs, [maxit] * 6,
                           coord_stride=3 * f.nheavy)
    s0 = ctx.score(dev["bids"][sg.FP64], dev["rid"], fx["gd"], [0] * 6, confs, 0)
    s1 = ctx.score(dev["bids"][sg.FP32], dev["rid"], fx["gd"], [0] * 6, confs, 1)
    for k in range(6):
        eo, co, xo = o.refine(m, confs[k], maxit)
        assert e[k] == eo and np.array_equal(cr[k], co) and np.array_equal(xr[k], xo)
        assert s0[k] == o.eval(m, confs[k])
        assert s1[k] == o.eval_intramolecular(m, confs[k])


@pytest.mark.gpu
def test_flex_dock_fp64_matches_oracle(fx):
    """the whole docking of the flexible configuration through the host
    facade (chains, merge, refinement, rescoring, modes) in FP64, with the
    chains cut to a tenth of their steps, against the oracle's dock"""
    frac, exh = 0.1, 4
    d = sg.Docking(fx["rigid"], fx["gd"], fp64=True, exhaustiveness=exh, seed=3, hydrophobic_scale=1.1,
                   step_fraction=frac)
    r = d.dock([fx["model"]])[0]
    e, c, ev = fx["o"].dock(fx["m"], sg.mc_params(), seed=3, exhaustiveness=exh, threads=exh, step_fraction=frac)
    print("flex dock energies", r["e"], flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)

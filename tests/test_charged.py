"""Charge-dependent scoring: ad4_scoring and dkoes_scoring
(src/lib/builtinscoring.cpp:86-90, :107-111) with four-component tables
(result_components, src/lib/result_components.h: type only, |q_a|, |q_b|
and q_a q_b dependent; weighted_terms::eval_fast,
src/lib/weighted_terms.cpp:52-62), the charge maps of the grid build
(src/lib/cache.cpp:172-187) and their use in grid::evaluate
(src/lib/grid.cpp:28-48).

CPU: the host's tables equal the oracle's, component by component; the
oracle's components equal the closed forms of electrostatic and
ad4_solvation (src/lib/everything.h:64-156) with the AutoDock solvation and
volume parameters; a one-atom ligand beside a one-atom receptor at a table
knot gives the closed-form energy on the direct back end, which fixes which
atom's charge meets which atom's volume (the table's (min type, max type)
orientation); under Vina, charges change nothing; the per-atom term table
carries the charged terms.  GPU: maps and charge maps, evaluations on both
back ends and the whole docking are bit-exact with the oracle in FP64, and
the FP32 evaluation meets the FP32 rule."""
import math

import numpy as np
import pytest

import pyoracle as po
import smina_gpu as sg
from common import discontinuity_flags, fp32_rule, gd_c0, heavy_types, random_confs

C_H, N_D, O_A = 2, 7, 13
SOLV = {C_H: -0.001430, N_D: -0.001620, O_A: -0.002510}
VOL = {C_H: 33.510300, N_D: 22.449300, O_A: 17.157300}
CHARGED = ["ad4_scoring", "dkoes_scoring"]


def with_charges(mol, seed, scale=0.6):
    """seeded partial charges in [-scale, scale], three decimals as PDBQT
    writes them"""
    rng = np.random.default_rng(seed)
    mol.set_charges(np.round(rng.uniform(-scale, scale, mol.natoms), 3))
    return mol


def receptor_q(seed=1):
    return with_charges(sg.Mol.receptor(seed, 1800, 8.0), 100 + seed)


def ligand_q(seed=21, nheavy=20, ntors=4):
    return with_charges(sg.Mol.ligand(seed, nheavy, ntors), 200 + seed)


def oracle_q(ligs, rec, scoring, gd=None):
    gd = gd or gd_c0()
    sf = sg.builtin_scoring(scoring)
    o = po.Oracle(sf, rec.xyz, rec.type, gd, receptor_charges=rec.charge)
    ids = [o.add_model(l.desc_ptr) for l in ligs]
    o.build_grids(heavy_types(*ligs))
    return o, ids, sf, gd


@pytest.fixture(scope="module")
def rec():
    return receptor_q()


@pytest.fixture(scope="module")
def lig():
    return ligand_q()


@pytest.mark.parametrize("name", CHARGED)
def test_tables_match_oracle(lig, rec, name):
    o, _, sf, _ = oracle_q([lig], rec, name)
    n, smooth, fast = sg.build_tables(sf)
    osm, ofa = o.tables()
    assert n == o.n and np.array_equal(smooth, osm) and np.array_equal(fast, ofa)
    k, qs, qf = sg.build_qtables(sf)
    ok, oqs, oqf = o.qtables()
    assert k == ok == 4 and np.array_equal(qs, oqs) and np.array_equal(qf, oqf)
    # the charge components are not all zero
    assert np.abs(qs).max() > 0 and np.abs(qf).max() > 0


def test_vina_has_one_component():
    k, qs, qf = sg.build_qtables(sg.vina_scoring())
    assert k == 1 and qs is None and qf is None


def ad4_parts(t1, t2, r, sigma=3.5, sq=0.01097):
    """ad4_solvation's four components for atoms (t1, t2) at r"""
    df = math.exp(-(r / (2 * sigma)) ** 2)
    return np.array([(SOLV[t1] * VOL[t2] + SOLV[t2] * VOL[t1]) * df, sq * VOL[t2] * df, sq * VOL[t1] * df, 0.0])


def test_component_closed_forms(lig, rec):
    """ad4_scoring: vdw(6-12) and non_dir_h_bond_lj in the first component
    only; ad4_solvation 0.1159 and electrostatic(i=1, cap 100) 0.1465"""
    o = oracle_q([lig], rec, "ad4_scoring")[0]
    for t1, t2, r in ((C_H, O_A, 3.7), (O_A, C_H, 3.7), (N_D, O_A, 2.9), (C_H, C_H, 6.5)):
        c = o.eval_components(t1, t2, r)
        want = 0.1159 * ad4_parts(t1, t2, r, 3.5)
        want[3] = 0.1465 * min(100.0, 1 / r)
        assert np.allclose(c[1:], want[1:], rtol=1e-13, atol=0), (t1, t2, r, c, want)
        # the first component: desolvation's type part plus the charge-independent terms
        assert c[0] - want[0] == pytest.approx(o.eval_terms(t1, t2, r) - want[0], rel=1e-12)
    # the |q_a| component follows the second atom's volume: swapping the types swaps it
    a, b = o.eval_components(C_H, O_A, 4.0), o.eval_components(O_A, C_H, 4.0)
    assert a[1] == b[2] and a[2] == b[1] and a[1] != a[2]
    # electrostatic's cap below r = 0.01
    assert o.eval_components(C_H, O_A, 0.005)[3] == pytest.approx(0.1465 * 100.0, rel=1e-14)
    # dkoes_scoring: ad4_solvation(sigma 3.6) 0.048934, no electrostatic
    od = oracle_q([lig], rec, "dkoes_scoring")[0]
    c = od.eval_components(C_H, O_A, 3.7)
    want = 0.048934 * ad4_parts(C_H, O_A, 3.7, 3.6)
    assert np.allclose(c[1:], want[1:], rtol=1e-13, atol=0) and c[3] == 0.0


def pdbqt_atom(n, name, x, q, ad, rec="LIG"):
    return f"ATOM  {n:5d} {name:<4s} {rec}     1    {x:8.3f}{0.0:8.3f}{0.0:8.3f}  0.00  0.00    {q:+6.3f} {ad:<2s}"


def one_atom_pair(q_lig, q_rec, t_lig="OA", t_rec="C", x=4.0):
    lig = sg.Mol.pdbqt_ligand("ROOT\n" + pdbqt_atom(1, "O1", 0.0, q_lig, t_lig) + "\nENDROOT\nTORSDOF 0\n")
    rec = sg.Mol.pdbqt_receptor(pdbqt_atom(1, "CA", x, q_rec, t_rec, rec="ALA") + "\n")
    return lig, rec


def closed_form_pair(t_lig, t_rec, q_lig, q_rec, r):
    """ad4_scoring's charged share for the pair at r: each atom's |charge|
    times the other atom's volume in the desolvation, the charges' product
    in the electrostatic term"""
    p = 0.1159 * ad4_parts(t_lig, t_rec, r)
    return p[1] * abs(q_lig) + p[2] * abs(q_rec) + 0.1465 * q_lig * q_rec * min(100.0, 1 / r)


def test_pdbqt_charges_are_read():
    lig, rec = one_atom_pair(-0.412, 0.275)
    assert lig.charge[0] == -0.412 and rec.charge[0] == 0.275


def test_one_pair_closed_form_on_the_direct_back_end():
    """ligand O_A (charge q_l) 4 A from a receptor C_H (charge q_r):
    r^2 = 16 is knot 512, so the direct back end's energy is the table's
    knot value exactly; the charged part (energy minus the energy with both
    charges 0) equals the closed form.  The ligand's type is the larger, so
    the table element is (C_H, O_A) with the atoms swapped."""
    box = sg.grid_dims((0.0, 0.0, 0.0), (12.0, 12.0, 12.0), 0.375)
    out = {}
    for q_l, q_r in ((-0.412, 0.275), (0.0, 0.0), (0.301, -0.5), (0.301, 0.0)):
        lig, rec = one_atom_pair(q_l, q_r)
        assert lig.type[0] == O_A and rec.type[0] == C_H
        o = po.Oracle(sg.builtin_scoring("ad4_scoring"), rec.xyz, rec.type, box, receptor_charges=rec.charge)
        mid = o.add_model(lig.desc_ptr)
        f = sg.Flat(lig)
        e, g, _ = o.eval_deriv(mid, f.initial_conf(), (1e30, 1e30, 1e30), backend=1, slope=0.0)
        out[(q_l, q_r)] = e
    base = out[(0.0, 0.0)]
    for (q_l, q_r), e in out.items():
        want = closed_form_pair(O_A, C_H, q_l, q_r, 4.0)
        assert e - base == pytest.approx(want, rel=1e-11, abs=1e-15), (q_l, q_r, e - base, want)


def test_vina_ignores_charges(lig):
    """Vina's tables have one component: the same ligand and receptor with
    and without charges score the same on both back ends"""
    plain_rec, plain_lig = sg.Mol.receptor(1, 1800, 8.0), sg.Mol.ligand(21, 20, 4)
    rec = receptor_q()
    sf = sg.vina_scoring()
    gd = gd_c0()
    res = []
    for r, l in ((plain_rec, plain_lig), (rec, lig)):
        o = po.Oracle(sf, r.xyz, r.type, gd, receptor_charges=r.charge)
        mid = o.add_model(l.desc_ptr)
        o.build_grids(heavy_types(l))
        c = random_confs(sg.Flat(l), gd, 3, seed0=77)
        res.append([o.eval_deriv(mid, x, (10.0, 1000.0, 10.0), backend=b)[0] for x in c for b in (0, 1)])
    assert res[0] == res[1]


def test_atom_terms_carry_the_charged_terms():
    """the per-atom table under ad4_scoring: charge-independent columns
    first, then ad4_solvation and electrostatic; the charged columns of a
    one-atom pair at 4 A equal the closed forms"""
    q_l, q_r = -0.412, 0.275
    lig, rec = one_atom_pair(q_l, q_r)
    sf = sg.builtin_scoring("ad4_scoring")
    text = lig.atom_terms(sf, rec, sg.Flat(lig).initial_conf())
    lines = text.strip().split("\n")
    head = lines[0].split(" ")
    assert head[3:] == ["vdw(i=6,_j=12,_s=0,_^=100,_c=8)", "non_dir_h_bond_lj(o=-0.7,_^=100,_c=8)",
                        "ad4_solvation(d-sigma=3.5,_s/q=0.01097,_c=8)", "electrostatic(i=1,_^=100,_c=8)"], head
    vals = [float(v) for v in lines[1][lines[1].index(">") + 1:].split()]
    p = ad4_parts(O_A, C_H, 4.0)
    assert vals[2] == pytest.approx(0.1159 * (p[0] + p[1] * abs(q_l) + p[2] * abs(q_r)), rel=1e-5)
    assert vals[3] == pytest.approx(0.1465 * q_l * q_r / 4.0, rel=1e-5)


def test_splines_refuse_charged_terms(rec):
    with pytest.raises(Exception, match="linear tables"):
        sg.Docking(rec, gd_c0(), scoring="ad4_scoring", approximation="spline")


# ------------------------------------------------------------------ device
@pytest.fixture(scope="module")
def dev(rec, lig):
    """both charged functions on the device: tables, the charged receptor,
    FP64 and FP32 maps (with charge maps) and the charged ligand"""
    out = {}
    for name in CHARGED:
        o, ids, sf, gd = oracle_q([lig], rec, name)
        ctx = sg.Context(0)
        n, smooth, fast = sg.build_tables(sf)
        _, qs, qf = sg.build_qtables(sf)
        ctx.upload_tables(n, smooth, fast, qsmooth=qs, qfast=qf)
        rid = ctx.upload_receptor(rec.xyz, rec.type, rec.charge)
        gids = {p: ctx.build_grids(rid, gd, heavy_types(lig), p) for p in (sg.FP64, sg.FP32)}
        f = sg.Flat(lig)
        bids = {p: ctx.upload_models([f], p) for p in (sg.FP64, sg.FP32)}
        out[name] = (o, ids, ctx, rid, gids, f, bids, gd)
    yield out
    for v in out.values():
        v[2].close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CHARGED)
def test_maps_fp64_bit_exact(dev, lig, name):
    o, ids, ctx, rid, gids, f, bids, gd = dev[name]
    for t in heavy_types(lig):
        assert np.array_equal(ctx.download_grid(gids[sg.FP64], t, gd), o.grid(t)), t
        q = ctx.download_grid(gids[sg.FP64], t, gd, charge=True)
        assert np.array_equal(q, o.charge_grid(t)), t
        assert np.abs(q).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", CHARGED)
def test_eval_fp64_bit_exact(dev, name):
    o, ids, ctx, rid, gids, f, bids, gd = dev[name]
    confs = random_confs(f, gd, 12, seed0=321)
    confs[6:, :3] = 0.0
    for backend in (0, 1):
        kw = dict(grid_id=gids[sg.FP64]) if backend == 0 else dict(grid_id=-1, receptor_id=rid, gd=gd, slope=10.0)
        e, g, x = ctx.eval_confs(bids[sg.FP64], [0] * len(confs), confs, f.change_size, f.natoms,
                                 (10.0, 1000.0, 10.0), **kw)
        for k, c in enumerate(confs):
            eo, go, xo = o.eval_deriv(ids[0], c, (10.0, 1000.0, 10.0), backend=backend,
                                      slope=1e6 if backend == 0 else 10.0)
            assert e[k] == eo and np.array_equal(g[k], go), (name, backend, k, e[k], eo)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CHARGED)
@pytest.mark.parametrize("backend", [0, 1])
def test_eval_fp32_rule(dev, rec, name, backend):
    """FP32 maps and tables against the FP64 oracle under the FP32 rule:
    99% within 1e-4, every evaluation away from a discontinuity within 1e-3"""
    o, ids, ctx, rid, gids, f, bids, gd = dev[name]
    confs = random_confs(f, gd, 200, seed0=900)
    v = (10.0, 10.0, 10.0)
    kw = dict(grid_id=gids[sg.FP32]) if backend == 0 else dict(grid_id=-1, receptor_id=rid, gd=gd, slope=10.0)
    e, g, x = ctx.eval_confs(bids[sg.FP32], [0] * len(confs), confs, f.change_size, f.natoms, v, **kw)
    ref = [o.eval_deriv(ids[0], c, v, backend=backend, slope=1e6 if backend == 0 else 10.0) for c in confs]
    flags = discontinuity_flags(f, gd, rec, [r[2] for r in ref], backend)
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    print(f"{name} backend {backend}: {within:.4f} within 1e-4, worst unflagged {worst:.2e}, {nflag} flagged")
    assert within >= 0.99 and worst <= 1e-3, (within, worst, nflag, k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ad4_scoring"])
def test_dock_fp64_matches_oracle(rec, lig, name):
    """the docking end to end in FP64 with charges on both sides: search on
    the maps with their charge maps, refinement on the direct term, final
    rescoring with the fast tables' four components"""
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=11, step_fraction=0.1, scoring=name)
    r = d.dock([lig])[0]
    o, ids, _, _ = oracle_q([lig], rec, name)
    e, c, ev = o.dock(ids[0], sg.mc_params(), seed=11, exhaustiveness=4, threads=4, step_fraction=0.1)
    print(f"{name}: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)


@pytest.mark.gpu
def test_dock_fp32_charged(rec, lig):
    """the served FP32 docking under ad4_scoring: poses found, FP64-rescored
    energies sorted, and each result's FP32-minus-FP64 error stated"""
    d = sg.Docking(rec, gd_c0(), exhaustiveness=8, seed=5, scoring="ad4_scoring")
    r = d.dock([lig])[0]
    es = r["e"]
    assert len(es) > 0 and all(np.isfinite(es)) and es == sorted(es)
    assert np.isfinite(r["fp64_error"]) and r["fp64_error"] >= 0

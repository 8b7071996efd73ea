"""Built-in scoring functions beside Vina (builtin_scoring,
src/lib/builtinscoring.cpp:65-113): vinardo (its own atom parameters,
src/lib/builtinscoring.cpp:6-62), dkoes_fast and dkoes_scoring_old (LJ
vdw, conf-independent num_tors_sqr and constant terms).

CPU: the host's tables equal the oracle's; the oracle's terms equal the
closed forms with vinardo's radii and hydrophobe flags; the final score
(conf-independent terms) equals the closed form on host and oracle.  GPU:
FP64 evaluation bit-exact and the whole docking equal to the oracle's."""
import math

import numpy as np
import pytest

import smina_gpu as sg
from common import gd_c0, heavy_types, oracle_for, random_confs, receptor

C_H, C_P, CA_H, CA_P, N_D, O_A, S_A = 2, 3, 4, 5, 7, 13, 15
NAMES = ["vina", "vinardo", "dkoes_fast", "dkoes_scoring_old"]


@pytest.fixture(scope="module")
def lig():
    return sg.Mol.ligand(21, 20, 4)


@pytest.mark.parametrize("name", NAMES)
def test_tables_match_oracle(lig, name):
    o, _, sf, _ = oracle_for([lig], scoring=name)
    n, smooth, fast = sg.build_tables(sf)
    osm, ofa = o.tables()
    assert n == o.n and np.array_equal(smooth, osm) and np.array_equal(fast, ofa)


def test_unknown_functions_are_refused():
    with pytest.raises(ValueError):
        sg.builtin_scoring("no_such_function")
    # the charge-dependent functions are built (their tests hold the charges)
    for name in ("dkoes_scoring", "ad4_scoring"):
        assert sg.builtin_scoring(name).nterms > 0


def test_vinardo_closed_forms(lig):
    """vinardo: gauss(0, 0.8) -0.045, repulsion 0.80, hydrophobic(0, 2.5)
    -0.035, non_dir_h_bond(-0.6, 0) -0.60; radii C 2.0 (aliphatic) / 1.9
    (aromatic), N 1.7, O 1.6; aromatic non-hydrophobe carbon and sulfur
    acceptor count as hydrophobes"""
    o = oracle_for([lig], scoring="vinardo")[0]
    g = lambda d: math.exp(-(d / 0.8) ** 2)
    # aliphatic carbons at their optimal distance 4.0: full gauss and hydrophobic
    assert o.eval_terms(C_H, C_H, 4.0) == pytest.approx(-0.045 - 0.035, abs=1e-15)
    # 0.5 A inside: repulsion d^2
    assert o.eval_terms(C_H, C_H, 3.5) == pytest.approx(-0.045 * g(-0.5) + 0.80 * 0.25 - 0.035, abs=1e-15)
    # aromatic non-hydrophobe carbon is a hydrophobe in vinardo (radius 1.9)
    assert o.eval_terms(CA_P, C_H, 3.9) == pytest.approx(-0.045 - 0.035, abs=1e-15)
    # h-bond donor/acceptor at 3.3 (0 at d = 0) and 2.7 (full at d = -0.6)
    assert o.eval_terms(N_D, O_A, 3.3) == pytest.approx(-0.045, abs=1e-15)
    assert o.eval_terms(N_D, O_A, 2.7) == pytest.approx(-0.045 * g(-0.6) + 0.80 * 0.36 - 0.60, abs=1e-15)
    # sulfur acceptor with an aliphatic carbon: hydrophobic contact at 4.0
    assert o.eval_terms(S_A, C_H, 4.0) == pytest.approx(-0.045 - 0.035, abs=1e-15)
    # the default parameters: the same aromatic carbon pair is no hydrophobe
    ov = oracle_for([lig], scoring="vina")[0]
    assert ov.eval_terms(CA_P, C_H, 3.8) == pytest.approx(-0.035579 + -0.005156 * math.exp(-2.25), abs=1e-15)


def test_dkoes_closed_forms(lig):
    """vdw(i=4, j=8, s=0, cap 100): depth -1 at the optimal distance, capped
    at 100 inside; non_dir_h_bond weights of dkoes_fast"""
    o = oracle_for([lig], scoring="dkoes_fast")[0]
    d0 = 3.8  # C-C, smina's default radii
    assert o.eval_terms(C_H, C_H, d0) == pytest.approx(-0.008962, abs=1e-15)
    r = 4.5
    lj = (-2 * d0 ** 4) / r ** 4 + d0 ** 8 / r ** 8
    assert o.eval_terms(C_H, C_H, r) == pytest.approx(0.008962 * lj, rel=1e-13)
    assert o.eval_terms(C_H, C_H, 1.0) == pytest.approx(0.008962 * 100, rel=1e-13)
    # N_D / O_A: vdw at the optimal distance 3.5 plus the full h-bond below 2.8
    assert o.eval_terms(N_D, O_A, 3.5) == pytest.approx(-0.008962, abs=1e-15)


@pytest.mark.parametrize("name,expect", [
    ("vina", lambda e, n: e / (1 + 0.1 * (5 * 0.05846 / 0.1) * n / 5)),
    ("vinardo", lambda e, n: e / (1 + 0.1 * (5 * 0.02 / 0.1) * n / 5)),
    ("dkoes_fast", lambda e, n: e + 0.1 * 0.285035 * n * n / 5 - 2.467357),
    ("dkoes_scoring_old", lambda e, n: e + 0.1 * 0.285035 * n * n / 5 - 2.585651),
])
def test_final_score_closed_forms(lig, name, expect):
    sf = sg.builtin_scoring(name)
    o = oracle_for([lig], scoring=name)[0]
    for e, n in ((-7.25, 3.0), (-11.5, 6.5), (2.0, 0.0)):
        host = sg.lib.sgpu_conf_independent(sf, n, e)
        assert host == o.conf_independent(n, e)
        assert host == pytest.approx(expect(e, n), rel=1e-14)


# ------------------------------------------------------------------ device
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["vinardo", "dkoes_fast"])
def test_eval_fp64_bit_exact(lig, name):
    o, ids, sf, gd = oracle_for([lig], scoring=name)
    rec = receptor()
    ctx = sg.Context(0)
    try:
        n, smooth, fast = sg.build_tables(sf)
        ctx.upload_tables(n, smooth, fast)
        rid = ctx.upload_receptor(rec.xyz, rec.type)
        gid = ctx.build_grids(rid, gd, heavy_types(lig), sg.FP64)
        f = sg.Flat(lig)
        bid = ctx.upload_models([f], sg.FP64)
        confs = random_confs(f, gd, 12, seed0=321)
        confs[6:, :3] = 0.0
        for backend in (0, 1):
            kw = dict(grid_id=gid) if backend == 0 else dict(grid_id=-1, receptor_id=rid, gd=gd, slope=10.0)
            e, g, x = ctx.eval_confs(bid, [0] * len(confs), confs, f.change_size, f.natoms, (10.0, 1000.0, 10.0),
                                     **kw)
            for k, c in enumerate(confs):
                eo, go, xo = o.eval_deriv(ids[0], c, (10.0, 1000.0, 10.0), backend=backend,
                                          slope=1e6 if backend == 0 else 10.0)
                assert e[k] == eo and np.array_equal(g[k], go), (name, backend, k)
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["vinardo", "dkoes_fast"])
def test_dock_fp64_matches_oracle(lig, name):
    """the docking end to end in FP64 under another built-in function,
    chains cut to a tenth of their steps on both sides"""
    d = sg.Docking(receptor(), gd_c0(), fp64=True, exhaustiveness=4, seed=11, step_fraction=0.1, scoring=name)
    r = d.dock([lig])[0]
    o, ids, _, _ = oracle_for([lig], scoring=name)
    e, c, ev = o.dock(ids[0], sg.mc_params(), seed=11, exhaustiveness=4, threads=4, step_fraction=0.1)
    print(f"{name}: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)

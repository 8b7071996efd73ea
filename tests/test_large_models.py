"""Models past one wavefront's width and at the layout's limits.

A ligand of 100 heavy atoms (109 atoms with polar hydrogens, 4,423 pairs,
25 frames, 30 degrees of freedom) takes every multi-pass path of the device
search: the atom loops and the FP32 prefix scans over two 64-atom chunks
(with the carry between them), pair loops of many batches and frames past a
row of 16 lanes.  FP64 evaluation, Monte Carlo and refinement bit-exact
against the oracle; FP32 within the served tolerance.  A model with more
pairs than the 13-bit pair field holds is refused with an error."""
import numpy as np
import pytest

import smina_gpu as sg
from common import discontinuity_flags, eval_confs_mix, fp32_rule, heavy_types, oracle_for, random_confs, receptor

pytestmark = pytest.mark.gpu

BIG = (5, 100, 24)


@pytest.fixture(scope="module")
def big():
    lig = sg.Mol.ligand(*BIG)
    o, ids, sf, gd = oracle_for([lig])
    rec = receptor()
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sf)
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(rec.xyz, rec.type)
    f = sg.Flat(lig)
    assert f.natoms > 64 and f.npairs > 4 * 4 * 64
    gids = {p: ctx.build_grids(rid, gd, heavy_types(lig), p) for p in (sg.FP64, sg.FP32)}
    bids = {p: ctx.upload_models([f], p) for p in (sg.FP64, sg.FP32)}
    yield dict(o=o, m=ids[0], gd=gd, ctx=ctx, rid=rid, f=f, gids=gids, bids=bids)
    ctx.close()


def evals(big, prec, backend, n=16):
    f, gd = big["f"], big["gd"]
    confs = eval_confs_mix(f, gd, n, seed=41)
    kw = dict(grid_id=big["gids"][prec]) if backend == 0 else dict(grid_id=-1, receptor_id=big["rid"], gd=gd,
                                                                   slope=10.0)
    e, g, x = big["ctx"].eval_confs(big["bids"][prec], [0] * n, confs, f.change_size, f.natoms, (10.0, 1000.0, 10.0),
                                    **kw)
    ref = [big["o"].eval_deriv(big["m"], c, (10.0, 1000.0, 10.0), backend=backend,
                               slope=1e6 if backend == 0 else 10.0) for c in confs]
    return e, g, x, ref


@pytest.mark.parametrize("backend", [0, 1])
def test_large_eval_fp64_bit_exact(big, backend):
    e, g, x, ref = evals(big, sg.FP64, backend)
    for k, (eo, go, xo) in enumerate(ref):
        assert e[k] == eo and np.array_equal(g[k], go) and np.array_equal(x[k], xo), k


@pytest.mark.parametrize("backend", [0, 1])
def test_large_eval_fp32_tolerance(big, backend):
    """the reference's FP32 rule: >= 99% within 1e-4 relative, every
    evaluation not flagged at a discontinuity within 1e-3"""
    e, g, x, ref = evals(big, sg.FP32, backend, n=100)
    flags = discontinuity_flags(big["f"], big["gd"], receptor(), [r[2] for r in ref], backend)
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    print(f"  {within:.4f} within 1e-4, worst unflagged {worst:.2e} (conf {k}), {nflag} flagged", flush=True)
    assert within >= 0.99 and worst <= 1e-3, (within, worst, nflag, k)
    for k, (eo, go, xo) in enumerate(ref):
        # coordinates: the FP32 frames place every atom within 1e-4 A
        assert np.max(np.abs(x[k] - xo)) < 1e-4, k


def test_large_mc_and_refine_fp64_bit_exact(big):
    ctx, o, m, f, gd = big["ctx"], big["o"], big["m"], big["f"], big["gd"]
    cs, ss = zip(*[f.make_chain(900 + k, gd) for k in range(3)])
    maxit = (25 + f.natoms) // 3
    cnt, e, c, x, ev = ctx.run_mc(big["bids"][sg.FP64], big["gids"][sg.FP64], sg.mc_params(), [0] * 3, np.array(cs),
                                  np.array(ss), [4] * 3, [maxit] * 3, coord_stride=3 * f.nheavy)
    for k in range(3):
        eo, co, xo, evo = o.mc_chain(m, sg.mc_params(), 4, maxit, cs[k], ss[k])
        assert cnt[k] == len(eo) and np.array_equal(e[k, :cnt[k]], eo) and ev[k] == evo
    confs = random_confs(f, gd, 3, seed0=77)
    confs[:, :3] = 0.25
    er, cr, xr = ctx.refine(big["bids"][sg.FP64], big["rid"], gd, [0] * 3, confs, [maxit] * 3,
                            coord_stride=3 * f.nheavy)
    for k in range(3):
        eo, co, xo = o.refine(m, confs[k], maxit)
        assert er[k] == eo and np.array_equal(cr[k], co), k


def test_model_past_the_pair_limit_is_refused(big):
    """150 heavy atoms: 8,904 pairs, past the 8,191 the 13-bit field holds"""
    f = sg.Flat(sg.Mol.ligand(7, 150, 20))
    assert f.npairs > 8191
    with pytest.raises(sg.SgpuError) as err:
        big["ctx"].upload_models([f], sg.FP32)
    assert "pairs" in str(err.value)

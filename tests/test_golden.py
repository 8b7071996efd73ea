"""The committed fixtures (tests/golden/c0_golden.npz, written by
tests/golden/make_golden.py) against the oracle on the CPU, and against the
FP64 device search on the GPU, both exactly."""
import os

import numpy as np
import pytest

import smina_gpu as sg
from common import receptor

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "c0_golden.npz")


def fixture():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixture_shapes():
    g = fixture()
    assert g["conf"].shape == (7, 13) and g["g_maps"].shape == (7, 12)
    assert g["grid_val"].shape == (len(g["grid_types"]), 256)
    # chains that save anything save at most 20 poses, best first
    for k, v in g.items():
        if k.startswith("mc_e_"):
            assert 1 <= len(v) <= 20 and np.all(np.diff(v) >= 0)


def test_oracle_reproduces_fixtures():
    import make_golden  # tests/golden on sys.path via conftest
    now = make_golden.compute()
    ref = fixture()
    assert sorted(now) == sorted(ref)
    for k in ref:
        assert np.array_equal(now[k], ref[k]), k


@pytest.mark.gpu
def test_fp64_device_reproduces_fixtures():
    import make_golden
    g = fixture()
    lig = sg.Mol.ligand(*make_golden.LIGAND)
    rec = receptor()
    from common import gd_c0
    gd = gd_c0()
    ctx = sg.Context(0)
    try:
        n, smooth, fast = sg.build_tables(sg.vina_scoring(1.0))
        ctx.upload_tables(n, smooth, fast)
        rid = ctx.upload_receptor(rec.xyz, rec.type)
        gid = ctx.build_grids(rid, gd, [int(t) for t in g["grid_types"]], sg.FP64)
        for t, vals in zip(g["grid_types"], g["grid_val"]):
            assert np.array_equal(ctx.download_grid(gid, int(t), gd)[g["grid_idx"]], vals)
        flat = sg.Flat(lig)
        bid = ctx.upload_models([flat], sg.FP64)
        confs = g["conf"]
        nc = len(confs)
        e, gr, _ = ctx.eval_confs(bid, [0] * nc, confs, flat.change_size, flat.natoms, (10.0, 10.0, 10.0),
                                  grid_id=gid)
        assert np.array_equal(e, g["e_maps"]) and np.array_equal(gr, g["g_maps"])
        e, gr, _ = ctx.eval_confs(bid, [0] * nc, confs, flat.change_size, flat.natoms, (1000.0, 1000.0, 1000.0),
                                  receptor_id=rid, gd=gd)
        assert np.array_equal(e, g["e_direct"]) and np.array_equal(gr, g["g_direct"])
        maxit = int(g["mc_maxiters"])
        for seed in make_golden.MC_SEEDS:
            for steps in make_golden.MC_STEPS:
                cnt, ee, cc, _, ev = ctx.run_mc(bid, gid, sg.mc_params(), [0], g[f"mc_start_{seed}"][None, :],
                                               g[f"mc_rng_{seed}"][None, :], [steps], [maxit],
                                               coord_stride=3 * flat.nheavy)
                ref = g[f"mc_e_{seed}_{steps}"]
                assert cnt[0] == len(ref) and np.array_equal(ee[0, :cnt[0]], ref)
                assert np.array_equal(cc[0, :cnt[0], :flat.conf_size], g[f"mc_c_{seed}_{steps}"])
                assert ev[0] == int(g[f"mc_evals_{seed}_{steps}"])
    finally:
        ctx.close()

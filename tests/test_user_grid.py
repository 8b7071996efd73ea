"""--user_grid: a user-supplied map added to the score.

The file (an AutoDock-style map: three skipped lines, SPACING, NELEMENTS,
CENTER, then one value per line) is read as setup_user_gd and grid::init
read it (src/main/main.cpp:772-810, src/lib/grid.cpp:77-108).  The map is
added where the reference adds it:
  - to every receptor map point at build time, evaluated at the point's
    integer indices taken as a location in A, with the cache slope
    (src/lib/cache.cpp:204-205);
  - to the direct receptor term of each heavy movable atom, value and
    derivative, with that evaluation's slope (src/lib/non_cache.cpp:180-184);
  - to the final rescoring at every ligand atom with slope 1000
    (src/lib/model.cpp:953-964).
The oracle's user-grid term is pinned against a numpy restatement of
grid::evaluate_aux (src/lib/grid.cpp:111-236) written from that text, bit
for bit, and its derivative against finite differences; the device against
the oracle, bit for bit in FP64."""
import numpy as np
import pytest

import pyoracle as po
import smina_gpu as sg
from common import gd_c0, heavy_types, ligand, oracle_for, receptor


def oracle_with_user_grid(ligs, ugd, data):
    """the oracle with the user grid set before its maps are built"""
    gd = gd_c0()
    o = po.Oracle(sg.vina_scoring(), receptor().xyz, receptor().type, gd)
    ids = [o.add_model(l.desc_ptr) for l in ligs]
    o.set_user_grid(ugd, data)
    o.build_grids(heavy_types(*ligs))
    return o, ids


def map_text(spacing, nelements, center, values, header=("GRID_PARAMETER_FILE r.gpf", "GRID_DATA_FILE r.maps.fld",
                                                          "MACROMOLECULE r.pdbqt")):
    lines = list(header) + [f"SPACING {spacing}", "NELEMENTS {} {} {}".format(*nelements),
                            "CENTER {} {} {}".format(*center)]
    lines += [repr(float(v)) for v in values]
    return "\n".join(lines) + "\n"


def big_user_grid(seed=3):
    """a 2 A user grid over [-12, 72]^3: it covers both the search box and
    the integer-index locations of the map build (0..60 A)"""
    nel = 41  # (41 + 1) x 2 A = 84 A
    n = nel + 1
    vals = np.random.default_rng(seed).uniform(-1.0, 0.5, n ** 3)
    return sg.read_user_grid(map_text(2.0, (nel, nel, nel), (29.0, 29.0, 29.0), vals))


def ug_value(ugd, data, loc, slope, v=1000.0):
    """grid::evaluate_aux without a derivative, restated from the text of
    src/lib/grid.cpp:111-236 (cap v, then the out-of-box penalty)"""
    dims = [ugd.n[i] + 1 for i in range(3)]
    s, miss, a = [0.0] * 3, [0.0] * 3, [0] * 3
    finv = []
    for i in range(3):
        dm1 = dims[i] - 1.0
        factor = dm1 / (ugd.end[i] - ugd.begin[i])
        finv.append(1 / factor)
        si = (loc[i] - ugd.begin[i]) * factor
        if si < 0:
            miss[i], a[i], si = -si, 0, 0.0
        elif si >= dm1:
            miss[i], a[i], si = si - dm1, dims[i] - 2, 1.0
        else:
            a[i] = int(si)
            si -= a[i]
        s[i] = si
    penalty = slope * (miss[0] * finv[0] + miss[1] * finv[1] + miss[2] * finv[2])

    def at(x, y, z):
        return data[x + dims[0] * (y + dims[1] * z)]
    x0, y0, z0 = a
    f000, f100, f010, f110 = at(x0, y0, z0), at(x0 + 1, y0, z0), at(x0, y0 + 1, z0), at(x0 + 1, y0 + 1, z0)
    f001, f101, f011, f111 = (at(x0, y0, z0 + 1), at(x0 + 1, y0, z0 + 1), at(x0, y0 + 1, z0 + 1),
                              at(x0 + 1, y0 + 1, z0 + 1))
    x, y, z = s
    mx, my, mz = 1 - x, 1 - y, 1 - z
    f = (f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz + f001 * mx * my * z +
         f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z)
    if f > 0:
        f *= v / (v + f)
    return f + penalty


# ------------------------------------------------------------------ CPU
def test_reader_known_answer():
    """NELEMENTS 2 2 2 at spacing 0.5: 3 intervals per axis (span (2 + 1) x
    0.5), the center moved by half a spacing, 27 values stored negated and
    scaled, the last plane of each axis zero"""
    vals = np.arange(1, 28, dtype=np.float64) / 10
    gd, data = sg.read_user_grid(map_text(0.5, (2, 2, 2), (1.0, 2.0, 3.0), vals), scaling=0.5)
    assert [gd.n[i] for i in range(3)] == [3, 3, 3]
    assert [gd.begin[i] for i in range(3)] == [0.5, 1.5, 2.5]
    assert [gd.end[i] for i in range(3)] == [2.0, 3.0, 4.0]
    assert data.shape == (64,)
    d = data.reshape(4, 4, 4)  # [z][y][x]
    assert np.array_equal(d[:3, :3, :3].reshape(-1), -(vals * 0.5))
    assert not d[3].any() and not d[:, 3].any() and not d[:, :, 3].any()
    # a short file: the missing values read as 0
    gd2, data2 = sg.read_user_grid(map_text(0.5, (2, 2, 2), (1.0, 2.0, 3.0), vals[:5]))
    assert np.array_equal(data2.reshape(4, 4, 4)[0, 0, :3], -vals[:3]) and not data2[8:].any()


def test_reader_refuses_a_bad_header():
    with pytest.raises(ValueError):
        sg.read_user_grid("a\nb\nc\nSPACING  0.5\nNELEMENTS 2 2 2\nCENTER 0 0 0\n1\n")


def test_user_grid_term_closed_forms():
    """the restated term: at a node inside the box the stored value itself,
    curled (x 1000 / (1000 + f)) when positive; between two nodes along x
    the linear blend; outside, the nearest face's value plus slope x the
    distance in A"""
    vals = np.array([-1.0, 2.0, -3.0] * 9)
    ugd, data = sg.read_user_grid(map_text(0.5, (2, 2, 2), (1.0, 2.0, 3.0), vals))
    assert data[0] == 1.0 and data[1] == -2.0
    assert ug_value(ugd, data, (0.5, 1.5, 2.5), 10.0) == 1.0 * (1000 / 1001)
    assert ug_value(ugd, data, (1.0, 1.5, 2.5), 10.0) == -2.0
    assert abs(ug_value(ugd, data, (0.75, 1.5, 2.5), 10.0) - (-0.5)) < 1e-15
    out = ug_value(ugd, data, (0.25, 1.5, 2.5), 10.0)
    assert abs(out - (1000 / 1001 + 10.0 * 0.25)) < 1e-12


def test_oracle_maps_carry_the_user_grid_at_integer_indices():
    """each map point gains the user grid evaluated at (x, y, z) taken as a
    location, with the cache slope: the oracle's maps with the user grid
    equal its maps without plus the restated term, bit for bit"""
    lig = ligand(3, 10, 0)
    ugd, data = big_user_grid()
    o0, _, _, gd = oracle_for([lig])
    o1, _ = oracle_with_user_grid([lig], ugd, data)
    t = heavy_types(lig)
    n = [gd.n[i] + 1 for i in range(3)]
    for typ in t[:1]:
        g0, g1 = o0.grid(typ), o1.grid(typ)
        for z in range(0, n[2], 7):
            for y in range(0, n[1], 5):
                for x in range(n[0]):
                    k = x + n[0] * (y + n[1] * z)
                    assert g1[k] == g0[k] + ug_value(ugd, data, (float(x), float(y), float(z)), 1e6), (x, y, z)


def test_oracle_user_grid_in_direct_term_and_rescoring():
    """the direct term's gradient with the user grid matches finite
    differences of its energy; the final rescoring (model::eval) adds the
    user grid at every ligand atom, slope 1000, in atom order"""
    lig = ligand(1, 24, 6)
    ugd, data = big_user_grid(5)
    o0, ids, _, gd = oracle_for([lig])
    o1, _ = oracle_with_user_grid([lig], ugd, data)
    f = sg.Flat(lig)
    c = f.initial_conf()
    v = (1000.0, 1000.0, 1000.0)
    e, g, x = o1.eval_deriv(ids[0], c, v, backend=1, slope=10.0)
    e0, g0, _ = o0.eval_deriv(ids[0], c, v, backend=1, slope=10.0)
    assert e != e0
    # the user grid's share of the position gradient against finite
    # differences of its share of the energy (the receptor and pair terms
    # cancel between the two oracles)
    h = 1e-5
    for k in range(3):
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        d = [o.eval_deriv(ids[0], cc, v, backend=1, slope=10.0)[0] for o in (o1, o0) for cc in (cp, cm)]
        fd = ((d[0] - d[1]) - (d[2] - d[3])) / (2 * h)
        an = g[k] - g0[k]
        assert abs(an) > 1e-2 and abs(fd - an) < 1e-3 * max(1.0, abs(an)), (k, fd, an)
    ev0 = o0.eval(ids[0], c)
    ev1 = o1.eval(ids[0], c)
    want = ev0
    for i in range(f.natoms):
        want += ug_value(ugd, data, tuple(x[3 * i:3 * i + 3]), 1000.0)
    assert ev1 == want


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def dev():
    ligs = [sg.Mol.ligand(1, 24, 6), sg.Mol.ligand(3, 10, 0)]
    ugd, data = big_user_grid(7)
    ou, ids_u = oracle_with_user_grid(ligs, ugd, data)
    sf, gd = sg.vina_scoring(), gd_c0()
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sf)
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(receptor().xyz, receptor().type)
    ctx.set_user_grid(ugd, data)
    types = heavy_types(*ligs)
    gids = {p: ctx.build_grids(rid, gd, types, p) for p in (sg.FP64, sg.FP32)}
    flats = [sg.Flat(l) for l in ligs]
    bids = {p: ctx.upload_models(flats, p) for p in (sg.FP64, sg.FP32)}
    yield dict(ligs=ligs, ugd=ugd, data=data, o=ou, ids=ids_u, gd=gd, ctx=ctx, rid=rid, gids=gids, flats=flats,
               bids=bids, types=types)
    ctx.close()


@pytest.mark.gpu
def test_user_grid_maps_fp64_bit_exact(dev):
    for t in dev["types"]:
        assert np.array_equal(dev["ctx"].download_grid(dev["gids"][sg.FP64], t, dev["gd"]), dev["o"].grid(t))


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1])
def test_user_grid_eval_fp64_bit_exact(dev, backend):
    from common import eval_confs_mix
    ctx, o = dev["ctx"], dev["o"]
    for li, f in enumerate(dev["flats"]):
        confs = eval_confs_mix(f, dev["gd"], 12, seed=40 + li)
        kw = dict(grid_id=dev["gids"][sg.FP64]) if backend == 0 else dict(grid_id=-1, receptor_id=dev["rid"],
                                                                         gd=dev["gd"], slope=10.0)
        e, g, x = ctx.eval_confs(dev["bids"][sg.FP64], [li] * len(confs), confs, f.change_size, f.natoms,
                                 (1000.0,) * 3, **kw)
        for k, c in enumerate(confs):
            eo, go, xo = o.eval_deriv(dev["ids"][li], c, (1000.0,) * 3, backend=backend,
                                      slope=1e6 if backend == 0 else 10.0)
            assert e[k] == eo and np.array_equal(g[k], go), (li, k, e[k], eo)
        s0 = ctx.score(dev["bids"][sg.FP64], dev["rid"], dev["gd"], [li] * len(confs), confs, 0)
        for k, c in enumerate(confs):
            assert s0[k] == o.eval(dev["ids"][li], c), (li, k)


@pytest.mark.gpu
def test_user_grid_dock_fp64_matches_oracle(dev):
    """a whole FP64 docking with --user_grid (maps, search, refinement and
    rescoring all carry it) equals the oracle's"""
    lig = dev["ligs"][0]
    d = sg.Docking(receptor(), gd_c0(), fp64=True, exhaustiveness=2, seed=5, step_fraction=0.05,
                   user_grid=(dev["ugd"], dev["data"]))
    r = d.dock([lig])[0]
    e, c, ev = dev["o"].dock(dev["ids"][0], sg.mc_params(), seed=5, exhaustiveness=2, threads=2, step_fraction=0.05)
    print(f"dock with a user grid: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)

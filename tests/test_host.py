"""The host side of the product against the oracle (CPU): scoring tables,
model flattening (pairs in smina's evaluation order, num_tors), the
initial conf and chain seeding, over ligands of the configurations' size
ranges; and the C ABI: libsgpu.so loads and exports every entry point the
headers declare."""
import os
import re
import subprocess

import numpy as np
import pytest

import smina_gpu as sg
from common import gd_c0, oracle_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LIGANDS = [(1, 24, 6), (2, 38, 10), (3, 10, 0), (4, 60, 15), (5, 13, 2), (6, 45, 7)]


@pytest.fixture(scope="module")
def setup():
    ligs = [sg.Mol.ligand(s, h, t) for s, h, t in LIGANDS]
    o, ids, sf, gd = oracle_for(ligs)
    return ligs, o, ids, sf, gd


def test_tables_match_oracle(setup):
    _, o, _, sf, _ = setup
    n, smooth, fast = sg.build_tables(sf)
    osm, ofa = o.tables()
    assert n == o.n and np.array_equal(smooth, osm) and np.array_equal(fast, ofa)
    # the custom scoring of the flexible configuration (hydrophobic x 1.1)
    sf2 = sg.vina_scoring(1.1)
    o2, _, _, _ = oracle_for([setup[0][0]], hydrophobic_scale=1.1)
    n2, s2, f2 = sg.build_tables(sf2)
    a2, b2 = o2.tables()
    assert np.array_equal(s2, a2) and np.array_equal(f2, b2) and not np.array_equal(s2, smooth)


@pytest.mark.parametrize("factor", [10.0, 32.0])
def test_splines_match_oracle(factor):
    """the product's precalculate_splines tables (the --minimize default)
    equal the oracle's bit for bit, for vina and vinardo"""
    lig = sg.Mol.ligand(1, 24, 6)
    for name in ("vina", "vinardo"):
        o, _, sf, _ = oracle_for([lig], scoring=name)
        nseg, coef, cut = sg.build_splines(sf, factor)
        ocoef, ocut = o.use_splines(factor)
        assert nseg == int(factor * 8) and np.array_equal(coef, ocoef) and np.array_equal(cut, ocut), name


def test_synthetic_ligands_have_requested_shape(setup):
    ligs = setup[0]
    for (seed, heavy, tors), lig in zip(LIGANDS, ligs):
        assert int(np.sum(lig.type[: lig.nmovable] > 1)) == heavy
        assert lig.nframes == tors + 1
        f = sg.Flat(lig)
        assert f.conf_size == 7 + tors and f.change_size == 6 + tors


def test_flattening_matches_oracle(setup):
    ligs, o, ids, _, gd = setup
    for lig, m in zip(ligs, ids):
        f = sg.Flat(lig)
        info = o.info(m)
        a, b, lst = o.pairs(m)
        assert f.npairs == info["npairs"] and f.nother == info["nother"]
        assert np.array_equal(a, f.pair_a) and np.array_equal(b, f.pair_b)
        assert np.all(lst == 0)  # a lone ligand: every pair is ligand-internal
        assert f.view.num_tors == int(o.num_tors(m))
        assert np.array_equal(f.initial_conf(), o.initial_conf(m))
        for seed in (1, 99, 123456):
            c1, s1 = f.make_chain(seed, gd)
            c2, s2 = o.randomize(m, seed)
            assert np.array_equal(c1, c2) and np.array_equal(s1, s2)


def test_pairs_exclude_close_and_rigid_atoms(setup):
    """every pair is two heavy atoms more than three bonds apart in
    different rigid pieces (src/lib/model.cpp:598-624)"""
    ligs = setup[0]
    for lig in ligs:
        f = sg.Flat(lig)
        d = lig.desc_ptr.contents
        nb = [[] for _ in range(lig.natoms)]
        for k in range(d.nbonds):
            a, b = d.bonds[2 * k], d.bonds[2 * k + 1]
            nb[a].append(b)
            nb[b].append(a)
        frame = np.zeros(lig.nmovable, int)
        for fr in range(d.nframes):
            frame[d.frame_begin[fr]:d.frame_end[fr]] = fr
        for a, b in zip(f.pair_a, f.pair_b):
            assert lig.type[a] > 1 and lig.type[b] > 1
            assert frame[a] != frame[b]
            # breadth-first bond distance > 3 (acyclic ligands)
            dist = {a: 0}
            q = [a]
            while q:
                x = q.pop(0)
                for y in nb[x]:
                    if y not in dist:
                        dist[y] = dist[x] + 1
                        q.append(y)
            assert dist.get(b, 99) > 3


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(sgpu_[a-z0-9_]+)\s*\(", text)) - {"sgpu_status"})


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "gpu-smina_amd", "libsgpu.so")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for h in ("sgpu.h", "sgpu_host.h"):
        names = declared(h)
        assert names, h
        missing = [n for n in names if n not in exported]
        assert not missing, missing
    # the binding resolved them too
    for name in sg.EXPORTED:
        assert hasattr(sg.lib, name)


def test_context_without_device_fails_cleanly():
    """no HIP device here: sgpu_create reports NO_DEVICE, never a fallback"""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(sg.SgpuError) as e:
        sg.Context(0)
    assert e.value.status == 5


@pytest.mark.parametrize("tasks,kw,expect", [
    # --fast (automatic split and fraction, early stopping while the device
    # is full) on a device holding 2,048 chains: the largest power of two
    # from 4 to 32 with tasks x split <= 4 x 2,048
    (8, {}, (32, 1.0, 0)),          # one ligand: the device is idle, full steps, no early stop
    (256, {}, (32, 1.0, 0)),        # 256 x 32 = 8,192: still 32
    (257, {}, (16, 0.5, 5)),        # one more: 16, half the steps, checkpoints round(10 x 0.5)
    (1024, {}, (8, 0.5, 5)),
    (8192, {}, (4, 0.5, 5)),        # the 1,024-ligand screen: the floor of 4
    (8192, dict(budget=0.0625), (4, 0.0625, 0)),  # a two-pass first pass: its budget, never stopped early
    # explicit settings (--gpu_split, --gpu_budget, --gpu_early_stop)
    (8, dict(chain_split=2), (2, 1.0, 10)),            # below 4: full steps; 10 checkpoints
    (8, dict(chain_split=1, step_fraction=1.0, early_stop=0), (1, 1.0, 0)),  # smina's search
    (8, dict(chain_split=1, step_fraction=0.3, early_stop=2), (1, 0.3, 3)),  # at least first + 1 = 3
    (8, dict(chain_split=32, early_stop=1), (32, 1.0, 0)),  # while the device is full only
    (8, dict(chain_split=32, early_stop=2), (32, 1.0, 10)),
])
def test_search_plan_known_answers(tasks, kw, expect):
    """the GPU-style search plan against the rules of auto_split,
    auto_fraction, stops and stop_for (src/gpu/gpu_docking.cpp:87-112,
    :167-193), worked out by hand"""
    assert sg.search_plan(tasks, 2048, **kw) == expect


def test_two_pass_cutoff_known_answers():
    """two-pass screening (src/main/main.cpp:1578-1586): the full search goes
    to ligands within the margin of the ceil(fast_top x n)-th lowest
    first-pass score"""
    scores = [-5.0 - 0.1 * k for k in range(30)]  # -5.0 .. -7.9
    # n = 30, fast_top 0.1: the 3rd lowest is -7.7; + 0.3
    assert sg.full_search_cutoff(scores) == pytest.approx(-7.4)
    # n = 31 after one more: ceil(3.1) = 4th lowest, -7.6
    assert sg.full_search_cutoff(scores + [-1.0]) == pytest.approx(-7.3)
    # fast_top 1: everything (the highest plus the margin)
    assert sg.full_search_cutoff(scores, 1.0, 0.0) == pytest.approx(-5.0)
    # a single score: itself plus the margin
    assert sg.full_search_cutoff([-6.0], 0.1, 0.3) == pytest.approx(-5.7)
    assert np.isnan(sg.full_search_cutoff([], 0.1, 0.3))

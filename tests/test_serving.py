"""The serving path (sgpu_server_*): receptors resident on the device, jobs
admitted as they arrive and batched continuously over concurrent launches.
Each job's result must be the same alone and in the mix: its chains are
seeded by the job alone and never interact with other jobs' chains."""
import numpy as np
import pytest

import smina_gpu as sg
from common import gd_c0, receptor

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    recs = [sg.Mol.receptor(11 + k, 1200 + 400 * k, 8.0) for k in range(2)]
    gds = [sg.grid_dims((0.0, 0.0, 0.0), (20.0, 20.0, 20.0)) for _ in recs]
    types = list(range(2, 14))  # the carbon, nitrogen and oxygen kinds (smina types 2-13)
    ligs = [sg.Mol.ligand(300 + k, 10 + 3 * k, min(k, 5)) for k in range(8)]
    exh = [2, 4, 2, 8, 4, 2, 4, 2]
    used = sorted({int(t) for l in ligs for t in l.type[:l.nmovable] if t > 1})
    assert set(used) <= set(types), used
    srv = sg.Server(workers=2, max_batch=3)
    ri = [srv.add_receptor(r, g, types) for r, g in zip(recs, gds)]
    # the mix: every job submitted at once, alternating receptors
    jobs = [srv.submit(ri[k % 2], l, exh[k]) for k, l in enumerate(ligs)]
    mixed = [srv.wait(j) for j in jobs]
    assert max(r["batch_size"] for r in mixed) > 1
    # alone: one job at a time
    for k, l in enumerate(ligs):
        alone = srv.wait(srv.submit(ri[k % 2], l, exh[k]))
        print(f"job {k}: batch {mixed[k]['batch_size']} e {mixed[k]['e'][:1]} latency {mixed[k]['latency_s']:.2f}s",
              flush=True)
        assert alone["batch_size"] == 1
        assert alone["e"] == mixed[k]["e"] and alone["evals"] == mixed[k]["evals"], k
    del srv


def test_server_matches_docking_facade():
    """a served job equals GpuDocking::dock_batch of the same ligand"""
    rec = sg.Mol.receptor(11, 1200, 8.0)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (20.0, 20.0, 20.0))
    lig = sg.Mol.ligand(77, 18, 3)
    types = sorted({int(t) for t in lig.type[:lig.nmovable] if t > 1})
    srv = sg.Server(workers=1)
    r = srv.wait(srv.submit(srv.add_receptor(rec, gd, types), lig, 4))
    d = sg.Docking(rec, gd, exhaustiveness=4)
    d.prepare([lig])
    ref = d.dock([lig])[0]
    assert r["e"] == ref["e"] and r["evals"] == ref["evals"]


@pytest.mark.gpu
def test_fast_serving_does_not_depend_on_the_batch():
    """--fast per job in the server (history check, a fixed split of 32): a
    job's modes are the same docked alone or beside others"""
    rec = receptor()
    gd = gd_c0()
    ligs = [sg.Mol.ligand(60 + k, 14 + 4 * k, 1 + k) for k in range(3)]
    types = list(range(2, 14))

    def run(batch):
        srv = sg.Server(workers=1, max_batch=8, fast=True)
        r = srv.add_receptor(rec, gd, types)
        jobs = [srv.submit(r, l, 4) for l in batch]
        return [srv.wait(j)["e"] for j in jobs]
    alone = run(ligs[:1])[0]
    mixed = run(ligs)[0]
    print("fast serving", alone[:3], mixed[:3], flush=True)
    assert len(alone) > 0 and alone == mixed


def test_a_failing_job_fails_alone():
    """a job that cannot be docked (a 300-heavy-atom ligand, past the
    layout's 255 atoms per model) beside good jobs in one launch: its wait
    raises with the reason, and the good jobs' results equal their results
    docked alone"""
    rec = sg.Mol.receptor(11, 1200, 8.0)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (20.0, 20.0, 20.0))
    good = [sg.Mol.ligand(400 + k, 12 + 2 * k, 2) for k in range(3)]
    bad = sg.Mol.ligand(500, 300, 4)
    srv = sg.Server(workers=1, max_batch=8)
    r = srv.add_receptor(rec, gd, list(range(2, 14)))
    jobs = [srv.submit(r, good[0], 2), srv.submit(r, bad, 2), srv.submit(r, good[1], 2), srv.submit(r, good[2], 2)]
    with pytest.raises(RuntimeError, match="atoms out of range"):
        srv.wait(jobs[1])
    mixed = [srv.wait(j) for j in (jobs[0], jobs[2], jobs[3])]
    assert max(m["batch_size"] for m in mixed) > 1
    for lig, m in zip(good, mixed):
        alone = srv.wait(srv.submit(r, lig, 2))
        assert len(m["e"]) > 0 and alone["e"] == m["e"] and alone["evals"] == m["evals"]
    del srv


def test_maps_resident_once_per_device():
    """every worker uses the store context's maps in place: the resident
    maps are the same bytes with one worker or four, and jobs docked by
    different workers equal a single worker's"""
    rec = sg.Mol.receptor(11, 1200, 8.0)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (20.0, 20.0, 20.0))
    types = list(range(2, 14))
    ligs = [sg.Mol.ligand(700 + k, 12 + 3 * k, 2) for k in range(4)]
    res = {}
    for workers in (1, 4):
        srv = sg.Server(workers=workers, max_batch=1)
        r = srv.add_receptor(rec, gd, types)
        res[workers] = ([srv.wait(j)["e"] for j in [srv.submit(r, l, 2) for l in ligs]], srv.memory())
        del srv
    print("memory", res[1][1], res[4][1], flush=True)
    assert res[1][1]["resident_maps"] == res[4][1]["resident_maps"] > 0
    assert res[1][0] == res[4][0]


def test_device_wide_budget_admits_by_memory():
    """one memory account for the device: with a budget of the maps plus
    one and a half jobs, launches in flight never hold more than that beside
    the maps, though three workers could run three at once (jobs wait for
    memory rather than overrun it); every job completes with the result it
    has with room to spare; a job that cannot fit beside the maps at all
    fails alone and the next job runs"""
    rec = sg.Mol.receptor(11, 1200, 8.0)
    gd = sg.grid_dims((0.0, 0.0, 0.0), (20.0, 20.0, 20.0))
    types = list(range(2, 14))
    ligs = [sg.Mol.ligand(720 + k, 14, 2) for k in range(6)]
    one = sg.Server(workers=1, max_batch=1)
    r1 = one.add_receptor(rec, gd, types)
    ref = [one.wait(j)["e"] for j in [one.submit(r1, l, 4) for l in ligs]]
    m1 = one.memory()
    single, maps = m1["peak_inflight"], m1["resident_maps"]
    del one
    srv = sg.Server(workers=3, max_batch=1, budget_gb=(maps + 1.5 * single) / 1e9)
    r = srv.add_receptor(rec, gd, types)
    got = [srv.wait(j)["e"] for j in [srv.submit(r, l, 4) for l in ligs]]
    m = srv.memory()
    print("one job", single, "maps", maps, "budgeted", m, flush=True)
    assert got == ref
    assert single <= m["peak_inflight"] <= 1.5 * single
    big = sg.Mol.ligand(730, 60, 10)
    jb = srv.submit(r, big, 64)
    jo = srv.submit(r, ligs[0], 4)
    with pytest.raises(RuntimeError, match="budget"):
        srv.wait(jb)
    assert srv.wait(jo)["e"] == ref[0]
    del srv


def test_served_fp64_job_matches_oracle_c4():
    """the serving configuration's sizes: 30 A boxes, receptors of 2,000 and
    6,000 heavy atoms, an 80-heavy-atom ligand with 20 torsions (and a small
    one) docked as FP64 server jobs at 1/50 of smina's steps; each equals the
    oracle's docking of the same ligand exactly"""
    from common import heavy_types, oracle_for
    gd = sg.grid_dims((0.0, 0.0, 0.0), (30.0, 30.0, 30.0))
    recs = [sg.Mol.receptor(100, 2000, 8.0), sg.Mol.receptor(101, 6000, 8.0)]
    ligs = [sg.Mol.ligand(5001, 80, 20), sg.Mol.ligand(5002, 16, 3)]
    types = list(range(2, 14))
    srv = sg.Server(workers=2, max_batch=4, fp64=True, step_fraction=0.02)
    ri = [srv.add_receptor(rc, gd, types) for rc in recs]
    jobs = [(k, i, srv.submit(ri[k], l, 2)) for k in range(2) for i, l in enumerate(ligs)]
    for k, i, j in jobs:
        r = srv.wait(j)
        o, ids, _, _ = oracle_for([ligs[i]], gd=gd, rec=recs[k])
        e, c, ev = o.dock(ids[0], sg.mc_params(), seed=1, exhaustiveness=2, threads=2, step_fraction=0.02)
        print(f"receptor {k} ligand {i}: served {r['e'][:3]} oracle {list(e[:3])}", flush=True)
        assert len(r["e"]) > 0 and r["e"] == list(e) and r["fp64_error"] is None
    del srv

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

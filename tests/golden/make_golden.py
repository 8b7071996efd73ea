"""Writes tests/golden/c0_golden.npz: fixtures of the docking search on the
single-ligand configuration (seeded 1,800-atom receptor with an 8 A pocket,
24-heavy-atom ligand with 6 torsions, 22.5 A box at 0.375 A), computed by
the CPU oracle (smina's FP64 search restated in C++):

  grid_*   receptor map values at 256 seeded points per ligand type;
  conf     seven confs (the parsed pose and six chain starts);
  e_maps, g_maps      model::eval_deriv on the maps, caps (10, 10, 10);
  e_direct, g_direct  the same on the receptor atoms (non_cache), caps 1000;
  mc_*     Monte Carlo chains (two seeds) at 1, 10 and 100 steps: saved
           energies, saved confs and evaluation counts.

The reference ships no data files for this path (its tests compare CPU
and GPU live), so these vectors are the oracle's own; the tests check that
the oracle and the FP64 device search both reproduce them exactly.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "gpu-smina_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import smina_gpu as sg  # noqa: E402
from common import heavy_types, oracle_for, random_confs  # noqa: E402

OUT = os.path.join(HERE, "c0_golden.npz")
LIGAND = (1, 24, 6)
MC_STEPS = (1, 10, 100)
MC_SEEDS = (3000, 3001)


def compute():
    """the fixture arrays, from the oracle"""
    lig = sg.Mol.ligand(*LIGAND)
    o, ids, sf, gd = oracle_for([lig])
    m = ids[0]
    flat = sg.Flat(lig)
    out = {}
    types = np.array(heavy_types(lig), np.int32)
    npts = (gd.n[0] + 1) * (gd.n[1] + 1) * (gd.n[2] + 1)
    idx = np.random.default_rng(11).integers(0, npts, 256).astype(np.int64)
    out["grid_types"] = types
    out["grid_idx"] = idx
    out["grid_val"] = np.array([o.grid(int(t))[idx] for t in types])
    confs = np.vstack([o.initial_conf(m)[None, :], random_confs(flat, gd, 6, seed0=700)])
    out["conf"] = confs
    for name, backend, v in (("maps", 0, (10.0, 10.0, 10.0)), ("direct", 1, (1000.0, 1000.0, 1000.0))):
        es, gs = [], []
        for c in confs:
            e, g, _ = o.eval_deriv(m, c, v, backend=backend)
            es.append(e)
            gs.append(g)
        out["e_" + name] = np.array(es)
        out["g_" + name] = np.array(gs)
    maxit = (25 + flat.natoms) // 3
    out["mc_maxiters"] = np.array(maxit)
    for seed in MC_SEEDS:
        c0, s0 = flat.make_chain(seed, gd)
        out[f"mc_start_{seed}"] = c0
        out[f"mc_rng_{seed}"] = s0
        for steps in MC_STEPS:
            e, c, x, ev = o.mc_chain(m, sg.mc_params(), steps, maxit, c0, s0)
            out[f"mc_e_{seed}_{steps}"] = e
            out[f"mc_c_{seed}_{steps}"] = c
            out[f"mc_evals_{seed}_{steps}"] = np.array(ev)
    return out


def main():
    np.savez_compressed(OUT, **compute())
    print("wrote", OUT)


if __name__ == "__main__":
    main()

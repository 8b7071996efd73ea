"""Shared builders for the tests: the seeded synthetic inputs of the
configurations (receptor of 1,800 heavy atoms with an 8 A pocket, ligands
with chosen heavy-atom and torsion counts) at sizes the oracle finishes in
seconds."""
import functools

import numpy as np

import pyoracle as po
import smina_gpu as sg

BOX_C0 = ((0.0, 0.0, 0.0), (22.5, 22.5, 22.5))


@functools.lru_cache(maxsize=None)
def receptor(seed=1, nheavy=1800, pocket=8.0):
    return sg.Mol.receptor(seed, nheavy, pocket)


@functools.lru_cache(maxsize=None)
def ligand(seed=1, nheavy=24, ntors=6):
    return sg.Mol.ligand(seed, nheavy, ntors)


@functools.lru_cache(maxsize=None)
def flex_model(seed=7, nflex=4, min_chi=2, max_chi=4):
    """(model, rigid receptor) of the flexible-receptor configuration: the
    24-atom ligand plus four side chains with 2-4 chi torsions each"""
    return sg.Mol.flex(receptor(), ligand(), seed, nflex, min_chi, max_chi)


def heavy_types(*mols):
    ts = set()
    for m in mols:
        ts.update(int(t) for t in m.type[: m.nmovable] if t > 1)
    return sorted(ts)


def gd_c0():
    return sg.grid_dims(*BOX_C0)


def oracle_for(ligands, gd=None, hydrophobic_scale=1.0, rec=None, scoring="vina"):
    rec = rec or receptor()
    gd = gd or gd_c0()
    sf = sg.builtin_scoring(scoring, hydrophobic_scale)
    o = po.Oracle(sf, rec.xyz, rec.type, gd)
    ids = [o.add_model(l.desc_ptr) for l in ligands]
    o.build_grids(heavy_types(*ligands))
    return o, ids, sf, gd


def random_confs(flat, gd, n, seed0=1000):
    """chain-start confs (random position in the box, random orientation
    and torsions) plus perturbations of the initial pose"""
    out = []
    for k in range(n):
        c, _ = flat.make_chain(seed0 + k, gd)
        out.append(c)
    return np.array(out)

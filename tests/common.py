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


def _q_mul(a, b):
    """Hamilton product of (w, x, y, z) quaternions"""
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                     a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                     a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])


def _inside_sphere(rng):
    while True:
        p = rng.uniform(-1.0, 1.0, 3)
        if p @ p < 1.0:
            return p


def eval_confs_mix(flat, gd, n, seed):
    """the conformations of the reference's FP32 evaluation test
    (src/gpu/tests/gpu_eval_test.cpp:96-125): even items fully random in the
    box (a chain start), odd items perturbed around the input pose: the
    position moved inside a unit sphere, the orientation turned by 0.3 x a
    vector inside the unit sphere, ligand torsions U(-0.5, 0.5) and
    side-chain torsions U(-pi, pi).  The synthetic models' input pose sits
    at the pocket (the box centre)."""
    rng = np.random.default_rng(seed)
    init = flat.initial_conf()
    v = flat.view
    tree = np.ctypeslib.as_array(v.frame_tree, (v.nframes,))
    conf_at = np.ctypeslib.as_array(v.frame_conf, (v.nframes,))
    parent = np.ctypeslib.as_array(v.frame_parent, (v.nframes,))
    lig_tors = [int(conf_at[f]) for f in range(v.nframes) if tree[f] < v.nligands and parent[f] >= 0]
    flex_tors = [int(conf_at[f]) for f in range(v.nframes) if tree[f] >= v.nligands]
    out = []
    for i in range(n):
        if i % 2 == 0:
            c, _ = flat.make_chain(seed * 1000 + i, gd)
        else:
            c = init.copy()
            c[:3] += _inside_sphere(rng)
            rot = 0.3 * _inside_sphere(rng)
            ang = np.linalg.norm(rot)
            dq = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * rot / ang]) if ang > 0 else \
                np.array([1.0, 0, 0, 0])
            q = _q_mul(dq, c[3:7])
            c[3:7] = q / np.linalg.norm(q)
            c[lig_tors] = rng.uniform(-0.5, 0.5, len(lig_tors))
            c[flex_tors] = rng.uniform(-np.pi, np.pi, len(flex_tors))
        out.append(c)
    return np.array(out)


from accuracy import discontinuity_flags, fp32_rule  # noqa: E402,F401  (shared with bench.py)

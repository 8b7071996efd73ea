"""SDF poses and the per-atom term table (--atom_terms) on the CPU.

--atom_terms writes, per pose, each ligand atom's weighted share of every
scoring term (result_info::setAtomValues / writeAtomValues,
src/lib/result_info.cpp:41-76, from terms::evale_robust,
src/lib/terms.cpp:235-307).  Pinned by a one-atom ligand beside a one-atom
receptor at the pair's optimal distance, where every Vina term has a closed
form, and by the whole table of a docked-size ligand summing, term by term
over atoms, to the oracle's intermolecular energy of the pose (the oracle
reads tabulated terms, so within the table's discretisation).  The SDF
writer (result_info::write's native branch, :139-163) is checked by reading
its molfile back: atoms at the pose, the model's bonds, the data items."""
import numpy as np

import smina_gpu as sg
from common import gd_c0, ligand, oracle_for, receptor


def atom(n, name, x, y, z, ad, rec="LIG"):
    return f"ATOM  {n:5d} {name:<4s} {rec}     1    {x:8.3f}{y:8.3f}{z:8.3f}  0.00  0.00    +0.000 {ad:<2s}"


VINA_NAMES = ["gauss(o=0,_w=0.5,_c=8)", "gauss(o=3,_w=2,_c=8)", "repulsion(o=0,_c=8)",
              "hydrophobic(g=0.5,_b=1.5,_c=8)", "non_dir_h_bond(g=-0.7,_b=0,_c=8)"]
VINA_W = [-0.035579, -0.005156, 0.840245, -0.035069, -0.587439]


def parse_table(text):
    lines = text.strip().split("\n")
    assert lines[-1] == "END"
    head = lines[0].split(" ")
    rows = []
    for ln in lines[1:-1]:
        pos = ln[ln.index("<") + 1:ln.index(">")]
        vals = [float(v) for v in ln[ln.index(">") + 1:].split()]
        rows.append((ln[:ln.index("<")].rstrip(), tuple(float(c) for c in pos.split(",")), vals))
    return head, rows


def test_atom_terms_closed_form():
    """two aliphatic carbons (C_H, X-Score radius 1.9) 3.8 A apart: surface
    distance 0, so gauss 1 = 1, gauss 2 = exp(-(3/2)^2), no repulsion, the
    hydrophobic term at its good end (1), no hydrogen bond; each weighted"""
    lig = sg.Mol.pdbqt_ligand("ROOT\n" + atom(1, "C1", 0.0, 0.0, 0.0, "C") + "\nENDROOT\nTORSDOF 0\n")
    rec = sg.Mol.pdbqt_receptor(atom(1, "CA", 3.8, 0.0, 0.0, "C", rec="ALA") + "\n")
    assert lig.type[0] == 2 and rec.type[0] == 2
    sf = sg.vina_scoring()
    f = sg.Flat(lig)
    head, rows = parse_table(lig.atom_terms(sf, rec, f.initial_conf()))
    assert head == ["atomid", "el", "pos"] + VINA_NAMES
    (label, pos, vals), = rows
    assert label == "    1 C1" and pos == (0.0, 0.0, 0.0)
    raw = [1.0, np.exp(-2.25), 0.0, 1.0, 0.0]
    want = [r * w for r, w in zip(raw, VINA_W)]
    assert np.allclose(vals, want, rtol=1e-5, atol=0), (vals, want)
    # beyond the cutoff: nothing
    far = sg.Mol.pdbqt_receptor(atom(1, "CA", 8.5, 0.0, 0.0, "C", rec="ALA") + "\n")
    assert not any(parse_table(lig.atom_terms(sf, far, f.initial_conf()))[1][0][2])


def test_atom_terms_sum_to_the_intermolecular_energy():
    """a rigid ligand (no torsions, so no intramolecular pairs) in the
    pocket: the table's weighted terms summed over atoms and terms equal the
    oracle's model::eval with the caps off, within the tables' resolution"""
    lig = ligand(3, 10, 0)
    rec = receptor()
    f = sg.Flat(lig)
    conf = f.initial_conf()
    head, rows = parse_table(lig.atom_terms(sg.vina_scoring(), rec, conf))
    assert len(rows) == lig.nmovable
    total = sum(sum(v) for _, _, v in rows)
    o, ids, _, _ = oracle_for([lig])
    e = o.eval(ids[0], conf, v=(1e30, 1e30, 1e30), slope=0.0)
    print("atom terms", total, "oracle", e)
    assert abs(total - e) <= 0.002 * abs(e) + 0.002, (total, e)
    # hydrogens carry no terms
    for (label, pos, vals), t in zip(rows, lig.type):
        if t in (0, 1):
            assert not any(vals)


def parse_sdf(text):
    mols = text.split("$$$$\n")
    assert mols[-1] == ""
    out = []
    for m in mols[:-1]:
        lines = m.split("\n")
        na, nb = int(lines[3][:3]), int(lines[3][3:6])
        assert lines[3].endswith("V2000")
        atoms = [(float(l[0:10]), float(l[10:20]), float(l[20:30]), l[31:34].strip()) for l in lines[4:4 + na]]
        bonds = [(int(l[0:3]), int(l[3:6]), int(l[6:9])) for l in lines[4 + na:4 + na + nb]]
        assert lines[4 + na + nb] == "M  END"
        data = {}
        rest = lines[5 + na + nb:]
        k = 0
        while k < len(rest):
            if rest[k].startswith("> <"):
                key = rest[k][3:-1]
                vals = []
                k += 1
                while k < len(rest) and rest[k] != "":
                    vals.append(rest[k])
                    k += 1
                data[key] = vals
            k += 1
        out.append((atoms, bonds, data))
    return out


def test_sdf_poses_read_back():
    lig = ligand(1, 24, 6)
    f = sg.Flat(lig)
    c0 = f.initial_conf()
    c1 = c0.copy()
    c1[:3] += (1.5, -2.0, 0.25)
    text = lig.write_sdf([c0, c1], [-7.123456, -6.5], rmsds=[0.0, -1.0],
                         atom_terms=(sg.vina_scoring(), receptor()))
    poses = parse_sdf(text)
    assert len(poses) == 2
    el = {0: "H", 1: "H", 2: "C", 3: "C", 6: "N", 7: "N", 8: "N", 9: "N", 10: "O", 11: "O", 12: "O", 13: "O"}
    (a0, b0, d0), (a1, b1, d1) = poses
    assert len(a0) == lig.nmovable and [a[3] for a in a0] == [el[int(t)] for t in lig.type[:lig.nmovable]]
    x0 = np.array([a[:3] for a in a0])
    assert np.allclose(x0, np.round(lig.xyz[:lig.nmovable], 4), atol=1.5e-4)
    x1 = np.array([a[:3] for a in a1])
    assert np.allclose(x1 - x0, (1.5, -2.0, 0.25), atol=2e-4)
    want_bonds = sorted(tuple(sorted((int(a) + 1, int(b) + 1))) for a, b in lig.bonds)
    assert sorted(tuple(sorted(b[:2])) for b in b0) == want_bonds and all(b[2] == 1 for b in b0)
    assert d0["minimizedAffinity"] == ["-7.12346"] and d0["minimizedRMSD"] == ["0.00000"]
    assert d1["minimizedAffinity"] == ["-6.50000"] and "minimizedRMSD" not in d1
    head = d0["atomic_interaction_terms"][0].split(" ")
    assert head[3:] == VINA_NAMES and d0["atomic_interaction_terms"][-1] == "END"

"""PDBQT input and output (CPU): smina's own parser and model set-up
(src/lib/parse_pdbqt.cpp:146-336, src/lib/model.cpp:448-566) and its pose
writer (src/lib/model.cpp:710-750, src/lib/result_info.cpp:146-153).

The reference ships no PDBQT files, so the expected models below are worked
out by hand from the parser's rules (atom order, frames, hydrogen-only
branches, bonds by covalent radii, X-Score typing from the bonds), and the
synthetic ligands are written out as PDBQT and read back: the read model
has the same atoms, types, frames, bonds and pairs as the one it was written
from, its energy moved only by the file's 3-decimal coordinates."""
import numpy as np
import pytest

import smina_gpu as sg
from common import oracle_for

C_H, C_P, O_A, O_DA, HD = 2, 3, 13, 12, 1


def atom(n, name, x, y, z, ad):
    return f"ATOM  {n:5d} {name:<4s} LIG     1    {x:8.3f}{y:8.3f}{z:8.3f}  0.00  0.00    +0.000 {ad:<2s}"


ETHANOL = "\n".join([
    "REMARK  ethanol: the hydroxyl branch moves only a hydrogen",
    "ROOT",
    atom(1, "C1", 0.0, 0.0, 0.0, "C"),
    atom(2, "C2", 1.52, 0.0, 0.0, "C"),
    "ENDROOT",
    "BRANCH   2   3",
    atom(3, "O1", 2.02, 1.35, 0.0, "OA"),
    atom(4, "H1", 2.98, 1.30, 0.0, "HD"),
    "ENDBRANCH   2   3",
    "TORSDOF 1",
]) + "\n"

# C1-C2 root; C2 -> C3 branch holding C4; C4 -> O5 branch holding C6
CHAIN = "\n".join([
    "ROOT",
    atom(1, "C1", 0.0, 0.0, 0.0, "C"),
    atom(2, "C2", 1.5, 0.0, 0.0, "C"),
    "ENDROOT",
    "BRANCH   2   3",
    atom(3, "C3", 2.0, 1.4, 0.0, "C"),
    atom(4, "C4", 3.5, 1.4, 0.0, "C"),
    "BRANCH   4   5",
    atom(5, "O5", 4.0, 2.8, 0.0, "OA"),
    atom(6, "C6", 5.4, 2.8, 0.0, "C"),
    "ENDBRANCH   4   5",
    "ENDBRANCH   2   3",
    "TORSDOF 2",
]) + "\n"


def test_hydrogen_only_branch_merges_and_types_from_bonds():
    m = sg.Mol.pdbqt_ligand(ETHANOL)
    # one rigid frame: the O-H branch joins the root, its atoms after the
    # root's own
    assert m.natoms == 4 and m.nframes == 1
    assert m.frames[0] == dict(parent=-1, begin=0, end=4, origin=0, axis=-1)
    assert sorted(map(tuple, m.bonds.tolist())) == [(0, 1), (1, 2), (2, 3)]
    # C1 carbon-only: hydrophobic; C2 next to O: polar; O with HD: donor-acceptor
    assert list(m.type) == [C_H, C_P, O_DA, HD]


def test_nested_branches_atom_order_and_frames():
    m = sg.Mol.pdbqt_ligand(CHAIN)
    # each frame's atoms in file order, each followed by the first atom of
    # its child branch: root [C1, C2, C3], [C4, O5], [C6]
    assert np.allclose(m.xyz[:, 0], [0.0, 1.5, 2.0, 3.5, 4.0, 5.4])
    assert m.frames == [dict(parent=-1, begin=0, end=3, origin=0, axis=-1),
                        dict(parent=0, begin=3, end=5, origin=2, axis=1),
                        dict(parent=1, begin=5, end=6, origin=4, axis=3)]
    assert sorted(map(tuple, m.bonds.tolist())) == [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)]
    # C4 and C6 touch the oxygen; the oxygen has no polar hydrogen
    assert list(m.type) == [C_H, C_H, C_H, C_P, O_A, C_P]


@pytest.mark.parametrize("text,why", [
    (CHAIN.replace("TORSDOF 2\n", ""), "TORSDOF"),
    (CHAIN.replace(" OA", " Xx"), "AutoDock type"),
    (CHAIN.replace("ENDBRANCH   4   5", "ENDBRANCH   4   6"), "inconsistent branch"),
    (CHAIN.replace("BRANCH   4   5", "BRANCH   9   5"), "no atom number 9"),
    (CHAIN.replace("ENDROOT\n", ""), "inappropriate tag"),
])
def test_syntax_errors_name_the_problem(text, why):
    with pytest.raises(ValueError, match=why):
        sg.Mol.pdbqt_ligand(text)


AD_NAME = {0: "H", 1: "HD", 2: "C", 3: "C", 4: "A", 5: "A", 6: "N", 7: "N", 8: "NA", 9: "NA", 10: "O", 11: "O",
           12: "OA", 13: "OA", 14: "S", 15: "SA", 16: "P", 17: "F", 18: "Cl", 19: "Br", 20: "I"}


def to_pdbqt(mol):
    """a ligand description written as PDBQT: each frame's atoms, then per
    child frame a BRANCH block opening with the child's origin atom"""
    fr = mol.frames
    children = {f: [c for c in range(len(fr)) if fr[c]["parent"] == f] for f in range(len(fr))}
    lines = []

    def rec(a):
        x, y, z = mol.xyz[a]
        lines.append(atom(a + 1, "X", x, y, z, AD_NAME[int(mol.type[a])]))

    def level(f):
        imm = {fr[c]["origin"] for c in children[f]}
        for a in range(fr[f]["begin"], fr[f]["end"]):
            if a not in imm:
                rec(a)

    def branches(f):
        for c in children[f]:
            lines.append(f"BRANCH {fr[c]['axis'] + 1:3d} {fr[c]['origin'] + 1:3d}")
            rec(fr[c]["origin"])
            level(c)
            branches(c)
            lines.append(f"ENDBRANCH {fr[c]['axis'] + 1:3d} {fr[c]['origin'] + 1:3d}")

    lines.append("ROOT")
    level(0)
    lines.append("ENDROOT")
    branches(0)
    lines.append(f"TORSDOF {len(fr) - 1}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed,heavy,tors", [(1, 24, 6), (2, 38, 10), (4, 60, 15), (3, 10, 0)])
def test_synthetic_ligands_read_back(seed, heavy, tors):
    """a synthetic ligand (built in smina's atom order) written as PDBQT and
    read back: the same atoms, types, frames and bonds, coordinates to the
    file's 3 decimals, the same pair list, and its energy moved only by that
    rounding; a second write and read changes nothing, energy included"""
    src = sg.Mol.ligand(seed, heavy, tors)
    back = sg.Mol.pdbqt_ligand(to_pdbqt(src))
    assert back.natoms == src.natoms and list(back.type) == list(src.type)
    assert back.frames == src.frames
    bond_set = lambda m: sorted((min(a, b), max(a, b)) for a, b in m.bonds.tolist())
    assert bond_set(back) == bond_set(src)
    assert np.abs(back.xyz - src.xyz).max() <= 5.0001e-4
    assert sg.Flat(back).npairs == sg.Flat(src).npairs
    again = sg.Mol.pdbqt_ligand(to_pdbqt(back))
    assert np.array_equal(again.xyz, back.xyz) and list(again.type) == list(back.type)
    assert again.frames == back.frames and np.array_equal(again.bonds, back.bonds)
    o, ids, _, _ = oracle_for([src, back, again])
    e = [o.eval(i, o.initial_conf(i)) for i in ids]
    assert abs(e[1] - e[0]) < max(0.05, 1e-3 * abs(e[0])) and e[2] == e[1]


def test_pose_writer_round_trip():
    """poses written back: at the parsed pose every atom keeps its
    coordinates; a translated conf moves every atom by the translation; a
    full turn of a torsion changes nothing; the header carries the energy"""
    m = sg.Mol.pdbqt_ligand(CHAIN)
    f = sg.Flat(m)
    c0 = f.initial_conf()
    moved = c0.copy()
    moved[:3] += (1.0, -2.0, 0.5)
    turned = c0.copy()
    turned[7] += 2 * np.pi
    out = m.write_pdbqt(np.array([c0, moved, turned]), [-7.25, -6.5, -6.0], [0.0, 1.5, 2.0])
    models = out.split("ENDMDL\n")[:-1]
    assert len(models) == 3
    assert models[0].startswith("MODEL 1\nREMARK minimizedAffinity -7.25\nREMARK minimizedRMSD 0\n")

    def coords(block):
        return np.array([[float(l[30:38]), float(l[38:46]), float(l[46:54])]
                         for l in block.splitlines() if l.startswith("ATOM")])
    inp = coords(CHAIN)
    assert np.allclose(coords(models[0]), inp, atol=1e-3)
    assert np.allclose(coords(models[1]), inp + (1.0, -2.0, 0.5), atol=1.5e-3)
    assert np.allclose(coords(models[2]), inp, atol=1e-3)
    # the non-atom lines pass through unchanged
    assert [l for l in models[0].splitlines()[3:] if not l.startswith("ATOM")] == \
        [l for l in CHAIN.splitlines() if not l.startswith("ATOM")]


def test_receptor_records_and_typing():
    """a rigid receptor: atom records in file order, other records ignored,
    types from the bonds (a carbon next to a nitrogen is polar, a nitrogen
    with a polar hydrogen donates)"""
    text = "\n".join([
        "REMARK  a peptide fragment",
        atom(1, "N", 0.0, 0.0, 0.0, "N"),
        atom(2, "H", -0.9, 0.4, 0.0, "HD"),
        atom(3, "CA", 1.45, 0.0, 0.0, "C"),
        atom(4, "CB", 2.0, 1.42, 0.0, "C"),
        atom(5, "CG", 3.5, 1.42, 0.0, "C"),
        "TER",
        "END",
    ]) + "\n"
    r = sg.Mol.pdbqt_receptor(text)
    assert r.natoms == 5 and r.nmovable == 0
    assert list(r.type) == [7, HD, C_P, C_H, C_H]  # N_D, HD, C_P, C_H, C_H


def receptor_pdbqt(mol):
    return "\n".join(atom(i + 1, "X", *mol.xyz[i], AD_NAME[int(mol.type[i])]) for i in range(mol.natoms)) + "\n"


@pytest.mark.gpu
def test_docking_from_pdbqt_files():
    """the path from files: a receptor and a ligand read from PDBQT, docked
    in FP64 through the host facade, equal the oracle's docking of the same
    parsed models; the modes written back as PDBQT carry the energies and
    put the top mode's heavy atoms where the docking left them"""
    from common import gd_c0
    rec = sg.Mol.pdbqt_receptor(receptor_pdbqt(sg.Mol.receptor(1)))
    lig = sg.Mol.pdbqt_ligand(to_pdbqt(sg.Mol.ligand(1, 24, 6)))
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=5, step_fraction=0.05)
    r = d.dock([lig], top_coords=True, confs=True)[0]
    o, ids, _, _ = oracle_for([lig], rec=rec)
    e, c, ev = o.dock(ids[0], sg.mc_params(), seed=5, exhaustiveness=4, threads=4, step_fraction=0.05)
    print(f"docked from PDBQT: {r['e']}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)
    text = lig.write_pdbqt(r["confs"], r["e"])
    assert text.count("MODEL ") == len(r["e"])
    first = text.split("ENDMDL\n")[0]
    assert f"REMARK minimizedAffinity {np.float32(r['e'][0]):.9g}" in first
    heavy = [i for i in range(lig.natoms) if lig.type[i] > 1]
    rows = [l for l in first.splitlines() if l.startswith("ATOM")]
    xyz = np.array([[float(l[30:38]), float(l[38:46]), float(l[46:54])] for l in rows])
    # atom records are in file order; the model's atom i sits on the line of
    # the same coordinates in the input
    src = np.array([[float(l[30:38]), float(l[38:46]), float(l[46:54])]
                    for l in to_pdbqt(sg.Mol.ligand(1, 24, 6)).splitlines() if l.startswith("ATOM")])
    order = [int(np.argmin(np.abs(src - lig.xyz[i]).sum(axis=1))) for i in heavy]
    assert np.allclose(xyz[order], r["coords"], atol=1.5e-3)


def flex_pdbqt(model, nlig_frames):
    """the side chains of a model description (trees after the ligand's)
    written as a flexible-residue file: per chain a residue whose ROOT holds
    the fixed anchor atom and its hydrogens, and whose branches follow the
    chain's frames (each opening with its origin atom)"""
    fr = model.frames
    children = {f: [c for c in range(len(fr)) if fr[c]["parent"] == f] for f in range(len(fr))}
    bonded = {}
    for a, b in model.bonds.tolist():
        bonded.setdefault(a, []).append(b)
        bonded.setdefault(b, []).append(a)
    lines = []

    def rec(a):
        x, y, z = model.xyz[a]
        lines.append(atom(a + 1, "X", x, y, z, AD_NAME[int(model.type[a])]))

    def body(f):
        imm = {fr[c]["origin"] for c in children[f]}
        for a in range(fr[f]["begin"], fr[f]["end"]):
            if a not in imm and a != fr[f]["origin"]:
                rec(a)
        for c in children[f]:
            lines.append(f"BRANCH {fr[c]['axis'] + 1:3d} {fr[c]['origin'] + 1:3d}")
            rec(fr[c]["origin"])
            body(c)
            lines.append(f"ENDBRANCH {fr[c]['axis'] + 1:3d} {fr[c]['origin'] + 1:3d}")

    for f in range(nlig_frames, len(fr)):
        if fr[f]["parent"] != -1:
            continue
        anchor = fr[f]["axis"]
        lines.append("BEGIN_RES SCH A   1")
        lines.append("ROOT")
        rec(anchor)
        for h in bonded.get(anchor, []):
            if model.type[h] == HD and h >= model.nmovable:
                rec(h)
        lines.append("ENDROOT")
        lines.append(f"BRANCH {anchor + 1:3d} {fr[f]['origin'] + 1:3d}")
        rec(fr[f]["origin"])
        body(f)
        lines.append(f"ENDBRANCH {anchor + 1:3d} {fr[f]['origin'] + 1:3d}")
        lines.append("END_RES SCH A   1")
    return "\n".join(lines) + "\n"


def test_flexible_residues_read_back():
    """side chains cut from the synthetic receptor, written as a
    flexible-residue file and read back beside the ligand: one tree per
    chain with the same torsions, each chain's first atom now fixed (smina
    keeps the first atom of a root branch in place), the same atoms.  The
    two models move alike: over several confs their energies differ by one
    constant, the pairs between fixed atoms that smina leaves out
    (model::initialize_pairs) and the synthetic layout counts, up to the
    file's rounding"""
    rec = sg.Mol.receptor(1, 4000, 9.0)
    lig = sg.Mol.ligand(7, 30, 6)
    model, rigid = sg.Mol.flex(rec, lig, 3, nflex=4)
    back_lig = sg.Mol.pdbqt_ligand(to_pdbqt(lig))
    back = sg.Mol.pdbqt_flex(back_lig, rigid, flex_pdbqt(model, lig.nframes))
    assert back.natoms == model.natoms and sorted(back.type) == sorted(model.type)
    assert back.nframes == model.nframes and back.nmovable == model.nmovable - 4
    roots = [f for f in back.frames[lig.nframes:] if f["parent"] == -1]
    assert len(roots) == 4 and all(f["origin"] >= back.nmovable and f["axis"] >= back.nmovable for f in roots)
    o, ids, _, _ = oracle_for([model, back], rec=rigid)
    c0 = o.initial_conf(ids[0])
    rng = np.random.default_rng(5)
    diffs = []
    for k in range(4):
        c = c0.copy()
        if k:
            c[7:] += rng.uniform(-0.6, 0.6, c.size - 7)
            c[:3] += rng.uniform(-0.5, 0.5, 3)
        # box slope 10 (refinement's): an atom outside the box moves the
        # penalty by the slope times the file's rounding
        diffs.append(o.eval(ids[0], c, slope=10.0) - o.eval(ids[1], c, slope=10.0))
    print("flex energy differences", diffs)
    assert max(diffs) - min(diffs) < 0.1


def test_receptor_and_writer_errors():
    """a multi-MODEL receptor and an unknown type are refused with the line;
    poses without rmsd say -1; a model that was not read from PDBQT cannot
    be written back"""
    with pytest.raises(ValueError, match="line 1: unexpected multi-MODEL"):
        sg.Mol.pdbqt_receptor("MODEL 1\n" + atom(1, "N", 0, 0, 0, "N") + "\n")
    with pytest.raises(ValueError, match="line 2:.*AutoDock type"):
        sg.Mol.pdbqt_receptor(atom(1, "N", 0, 0, 0, "N") + "\n" + atom(2, "Q", 1, 0, 0, "Qq") + "\n")
    m = sg.Mol.pdbqt_ligand(CHAIN)
    c0 = sg.Flat(m).initial_conf()
    assert "REMARK minimizedRMSD -1\n" in m.write_pdbqt(c0[None], [-5.0])
    with pytest.raises(ValueError):
        sg.Mol.ligand(1, 24, 6).write_pdbqt(c0[None], [-5.0])

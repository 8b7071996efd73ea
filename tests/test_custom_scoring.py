"""--custom_scoring files (sgpu_parse_custom_scoring; custom_terms::add and
add_terms_from_file, src/lib/custom_terms.cpp:17-71, with the term forms of
src/lib/everything.h): the built-in functions written as files give the
built-ins' own terms and tables, charge-dependent terms follow the others
whatever the file order, weights scale, and terms this build does not
evaluate are refused with their line."""
import numpy as np
import pytest

import smina_gpu as sg

# src/lib/builtinscoring.cpp:65-113 written in the file format
VINA = """# the default Vina function
-0.035579    gauss(o=0,_w=0.5,_c=8)
-0.005156    gauss(o=3,_w=2,_c=8)
 0.840245    repulsion(o=0,_c=8)
-0.035069    hydrophobic(g=0.5,_b=1.5,_c=8)
-0.587439    non_dir_h_bond(g=-0.7,_b=0,_c=8)
 1.923       num_tors_div
"""
# ad4_scoring with the charge-dependent terms first: they still come last
AD4 = """0.1159  ad4_solvation(d-sigma=3.5,_s/q=0.01097,_c=8)   the charge terms first in the file
0.1465  electrostatic(i=1,_^=100,_c=8)
0.1560  vdw(i=6,_j=12,_s=0,_^=100,_c=8)

0.2744  num_tors_add
-0.0974 non_dir_h_bond_lj(o=-0.7,_^=100,_c=8)
"""


def terms_of(sf):
    return ([(sf.terms[i].kind, tuple(sf.terms[i].p), sf.terms[i].weight) for i in range(sf.nterms)],
            [(sf.ci[i].kind, sf.ci[i].weight) for i in range(sf.nci)])


def test_vina_file_is_the_builtin():
    sf = sg.custom_scoring(VINA.replace("1.923      ", repr(5 * 0.05846 / 0.1 - 1)))
    ref = sg.builtin_scoring("vina")
    assert terms_of(sf) == terms_of(ref)
    n1, s1, f1 = sg.build_tables(sf)
    n2, s2, f2 = sg.build_tables(ref)
    assert n1 == n2 and np.array_equal(s1, s2) and np.array_equal(f1, f2)


def test_charge_dependent_terms_follow_and_tables_match():
    sf = sg.custom_scoring(AD4)
    ref = sg.builtin_scoring("ad4_scoring")
    t, ci = terms_of(sf)
    assert [k for k, _, _ in t] == [sg.TERM_VDW, sg.TERM_NON_DIR_H_BOND_LJ, sg.TERM_AD4_SOLVATION,
                                    sg.TERM_ELECTROSTATIC]
    assert (t, ci) == terms_of(ref)
    for build in (sg.build_tables, sg.build_qtables):
        a, b = build(sf), build(ref)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_scale_comments_and_trailing_text():
    sf = sg.custom_scoring("# comment\n\n2.0 gauss(o=1,_w=0.3,_c=8) trailing words\n-1 constant_term\n", scale=0.25)
    t, ci = terms_of(sf)
    assert t == [(sg.TERM_GAUSS, (1.0, 0.3, 8.0, 0.0, 0.0), 0.5)]
    assert ci == [(sg.CI_CONSTANT, -0.25)]


@pytest.mark.parametrize("line,why", [
    ("1.0 ligand_length", "does not evaluate"),
    ("1.0 atom_type_gaussian(t1=C,t2=N,o=0,_w=1,_c=8)", "does not evaluate"),
    ("1.0 gauss(o=0,_c=8)", "unknown term"),
    ("1.0 gaussian(o=0,_w=1,_c=8)", "unknown term"),
    ("1.0 gauss(o=zero,_w=1,_c=8)", "could not convert"),
    ("1.0 electrostatic(i=3,_^=100,_c=8)", "1 or 2"),
    ("1.0 vdw(i=4,_j=12,_s=0,_^=100,_c=8)", "LJ exponents"),
])
def test_refusals_name_the_line(line, why):
    with pytest.raises(ValueError) as e:
        sg.custom_scoring("-0.0356 gauss(o=0,_w=0.5,_c=8)\n" + line + "\n")
    assert "line 2" in str(e.value) and why in str(e.value)


CUSTOM = """# pair terms the built-in functions leave out, and two conf-independent ones
-0.030   gauss(o=0,_w=0.5,_c=8)
-0.004   gauss(o=2,_w=0.9,_c=8)
 0.800   repulsion(o=0,_c=8)
-0.030   hydrophobic(g=0.5,_b=1.5,_c=8)
-0.020   non_hydrophobic(g=0.5,_b=1.5,_c=8)
-0.500   non_dir_h_bond(g=-0.7,_b=0,_c=8)
 0.050   non_dir_anti_h_bond_quadratic(o=0.5,_c=8)
 0.030   donor_donor_quadratic(o=0,_c=8)
 0.030   acceptor_acceptor_quadratic(o=0,_c=8)
 0.100   num_tors_sqrt
-0.300   constant_term
"""


@pytest.mark.gpu
def test_dock_fp64_custom_scoring_matches_oracle():
    """a docking with a --custom_scoring function through the host facade,
    FP64 at a tenth of the steps, against the oracle's dock with the parsed
    function: tables, maps, search, refinement and final score"""
    import pyoracle as po
    from common import gd_c0, heavy_types, receptor
    rec, lig = receptor(), sg.Mol.ligand(1, 24, 6)
    d = sg.Docking(rec, gd_c0(), fp64=True, exhaustiveness=4, seed=17, step_fraction=0.1, custom_scoring=CUSTOM)
    r = d.dock([lig])[0]
    sf = sg.custom_scoring(CUSTOM)
    o = po.Oracle(sf, rec.xyz, rec.type, gd_c0())
    m = o.add_model(lig.desc_ptr)
    o.build_grids(heavy_types(lig))
    e, c, ev = o.dock(m, sg.mc_params(), seed=17, exhaustiveness=4, threads=8, step_fraction=0.1)
    print(f"dock with a custom scoring file: {r['e'][:3]}", flush=True)
    assert len(r["e"]) > 0 and r["e"] == list(e)


def test_names_written_by_atom_terms_read_back():
    """the term names the --atom_terms table writes in its header (smina's
    parameterized forms) are read back into the same terms"""
    from common import receptor
    text = CUSTOM + "0.0489 ad4_solvation(d-sigma=3.6,_s/q=0.01097,_c=8)\n0.1465 electrostatic(i=2,_^=100,_c=8)\n" \
        "0.0089 vdw(i=4,_j=8,_s=0,_^=100,_c=8)\n-0.0974 non_dir_h_bond_lj(o=-0.7,_^=100,_c=8)\n"
    sf = sg.custom_scoring(text)
    lig = sg.Mol.ligand(1, 24, 6)
    f = sg.Flat(lig)
    head = lig.atom_terms(sf, receptor(), f.initial_conf()).strip().split("\n")[0].split()
    names = [h for h in head if "(" in h]
    assert len(names) == sf.nterms
    back = sg.custom_scoring("".join(f"1.0 {n}\n" for n in names))
    got = sorted((back.terms[i].kind, tuple(back.terms[i].p)) for i in range(back.nterms))
    want = sorted((sf.terms[i].kind, tuple(sf.terms[i].p)) for i in range(sf.nterms))
    assert got == want

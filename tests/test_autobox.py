"""--autobox_ligand / --autobox_add (sgpu_autobox; setup_autobox,
src/main/main.cpp:648-687, and the box construction of :800-809), against
boxes worked out by hand from the atoms' coordinates."""
import numpy as np
import pytest

import smina_gpu as sg


def atom(n, name, x, y, z, ad):
    return f"ATOM  {n:5d} {name:<4s} LIG     1    {x:8.3f}{y:8.3f}{z:8.3f}  0.00  0.00    +0.000 {ad:<2s}"


LIG = "\n".join(["ROOT", atom(1, "C1", 1.0, -2.0, 0.5, "C"), atom(2, "N1", 3.5, 0.0, -1.5, "N"),
                 atom(3, "O1", 2.0, 1.25, 4.0, "OA"), "ENDROOT", "TORSDOF 0", ""])


def test_box_of_one_ligand():
    gd = sg.autobox([sg.Mol.pdbqt_ligand(LIG)])
    # x in [1, 3.5]: centre 2.25, extent 2.5 + 8 = 10.5 -> 28 intervals of 0.375 (10.5 exactly)
    # y in [-2, 1.25]: centre -0.375, 3.25 + 8 = 11.25 -> 30 intervals
    # z in [-1.5, 4]: centre 1.25, 5.5 + 8 = 13.5 -> 36 intervals
    for d, (c, n) in enumerate([(2.25, 28), (-0.375, 30), (1.25, 36)]):
        assert gd.n[d] == n
        assert abs((gd.begin[d] + gd.end[d]) / 2 - c) < 1e-12
        assert abs(gd.end[d] - gd.begin[d] - 0.375 * n) < 1e-12


def test_size_override_add_and_several_molecules():
    other = LIG.replace("   1.000  -2.000   0.500", "  -3.000  -2.000   0.500")
    gd = sg.autobox([sg.Mol.pdbqt_ligand(LIG), sg.Mol.pdbqt_ligand(other)], add=1.0, size=(0.0, 20.0, 0.0))
    # x over both: [-3, 3.5], centre 0.25, 6.5 + 2 = 8.5 -> ceil(22.67) = 23 intervals
    assert gd.n[0] == 23 and abs((gd.begin[0] + gd.end[0]) / 2 - 0.25) < 1e-12
    # y: the given size 20 -> ceil(53.33) = 54 intervals
    assert gd.n[1] == 54 and abs(gd.end[1] - gd.begin[1] - 54 * 0.375) < 1e-12
    # the same box as grid_dims of the same centre and size
    ref = sg.grid_dims((0.25, -0.375, 1.25), (8.5, 20.0, 7.5))
    assert [gd.n[d] for d in range(3)] == [ref.n[d] for d in range(3)]
    assert np.allclose([gd.begin[d] for d in range(3)], [ref.begin[d] for d in range(3)], rtol=0, atol=1e-12)


def test_bad_arguments_are_refused():
    with pytest.raises(ValueError):
        sg.autobox([])
    with pytest.raises(ValueError):
        sg.autobox([sg.Mol.pdbqt_ligand(LIG)], spacing=0.0)

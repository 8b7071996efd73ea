"""The FP32 evaluation rule and its discontinuity flags
(gpu-smina_amd/accuracy.py), against hand-placed atoms and hand-made
errors: the reference's rule (src/gpu/tests/gpu_eval_test.cpp:55-92) asks
99% of evaluations within 1e-4 and every evaluation away from a
discontinuity (:186-233: a map cell face, the cutoff) within 10 x 1e-4."""
import numpy as np

import smina_gpu as sg
from common import discontinuity_flags, fp32_rule


def atom(n, name, x, y, z, ad, rec="LIG"):
    return f"ATOM  {n:5d} {name:<4s} {rec}     1    {x:8.3f}{y:8.3f}{z:8.3f}  0.00  0.00    +0.000 {ad:<2s}"


def setup():
    lig = sg.Mol.pdbqt_ligand("ROOT\n" + atom(1, "C1", 0.0, 0.0, 0.0, "C") + "\nENDROOT\nTORSDOF 0\n")
    rec = sg.Mol.pdbqt_receptor(atom(1, "CA", 0.0, 0.0, 0.0, "C", rec="ALA") + "\n")
    gd = sg.grid_dims((0.0, 0.0, 0.0), (12.0, 12.0, 12.0), 0.375)  # box -6..6, 32 intervals
    return sg.Flat(lig), rec, gd


def test_cell_faces_cutoff_and_box_edges_are_flagged():
    flat, rec, gd = setup()
    spacing = (gd.end[0] - gd.begin[0]) / gd.n[0]
    face = gd.begin[0] + 5 * spacing            # on a cell face in x
    mid = gd.begin[0] + 5.37 * spacing          # well inside a cell

    def flags(p, backend):
        return bool(discontinuity_flags(flat, gd, rec, [np.array([p])], backend)[0])

    # the receptor atom at the origin is 6.9 A away: clear of the 8 A cutoff
    inner = [mid, gd.begin[1] + 3.41 * spacing, gd.begin[2] + 7.63 * spacing]
    assert not flags(inner, 0) and not flags(inner, 1)
    # a cell face: the maps' trilinear interpolation switches cells there
    on_face = [face, inner[1], inner[2]]
    assert flags(on_face, 0)
    assert not flags(on_face, 1)  # the direct back end has no cells
    assert flags([face + 0.5e-4 * spacing, inner[1], inner[2]], 0)
    assert not flags([face + 5e-3 * spacing, inner[1], inner[2]], 0)
    # the box edge, where the out-of-box penalty starts
    assert flags([gd.end[0], inner[1], inner[2]], 0)
    # a receptor atom at the cutoff: r^2 within 2e-4 of 64 on the direct back
    # end (no cell faces there), measured from the box-clamped position
    def rec_at(x):
        return sg.Mol.pdbqt_receptor(atom(1, "CA", x, 0.1, 0.1, "C", rec="ALA") + "\n")

    def direct(p, r):
        return bool(discontinuity_flags(flat, gd, r, [np.array([p])], 1)[0])

    assert direct([0.0, 0.1, 0.1], rec_at(8.0))            # r^2 = 64
    assert direct([0.0, 0.1, 0.1], rec_at(8.0 + 1e-5))     # r^2 = 64.00016
    assert not direct([0.05, 0.1, 0.1], rec_at(8.0))       # r^2 = 63.2
    # outside the box the clamped position counts: x = -7 clamps to -6,
    # 8 A from a receptor atom at x = 2 (9 A from the atom itself)
    assert direct([-7.0, 0.1, 0.1], rec_at(2.0))
    assert not direct([-5.5, 0.1, 0.1], rec_at(2.0))


def test_rule_counts_and_worst_unflagged():
    n = 200
    ref = [(-5.0, np.array([1.0, -2.0, 0.5]), None) for _ in range(n)]
    e = np.full(n, -5.0)
    g = np.array([r[1].copy() for r in ref])
    flags = np.zeros(n, bool)
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    assert within == 1.0 and worst == 0.0 and nflag == 0
    # one unflagged evaluation 5e-4 off in energy (relative to max(1, |E|) = 5)
    e[7] = -5.0 + 5 * 5e-4
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    assert within == (n - 1) / n and abs(worst - 5e-4) < 1e-12 and k == 7
    # a large error at a flagged evaluation counts against the 99% only
    g[9, 1] += 10.0
    flags[9] = True
    within, worst, nflag, k = fp32_rule(e, g, ref, flags)
    assert within == (n - 2) / n and abs(worst - 5e-4) < 1e-12 and nflag == 1 and k == 7
    # the gradient error is relative to max(1, max |g|) = 2: 6e-4 absolute
    # is 3e-4, below the energy's 5e-4; 1.6e-3 absolute is 8e-4 and the worst
    g2 = g.copy()
    g2[3, 0] += 6e-4
    _, worst2, _, k2 = fp32_rule(e, g2, ref, flags)
    assert abs(worst2 - 5e-4) < 1e-12 and k2 == 7
    g2[3, 0] += 1e-3
    _, worst2, _, k2 = fp32_rule(e, g2, ref, flags)
    assert abs(worst2 - 8e-4) < 1e-12 and k2 == 3

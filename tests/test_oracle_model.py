"""Pins of the oracle's model and scoring (CPU): closed-form term values,
table construction invariants, tree invariants, finite-difference
gradients, brute-force receptor maps and BFGS on a convex quadratic."""
import numpy as np
import pytest

import pyoracle as po
import smina_gpu as sg
from common import gd_c0, ligand, oracle_for, receptor, random_confs

W = dict(g1=-0.035579, g2=-0.005156, rep=0.840245, hyd=-0.035069, hb=-0.587439)
C_H, C_P, N_D, N_DA, O_A, O_DA = 2, 3, 7, 8, 13, 12


@pytest.fixture(scope="module")
def c0():
    lig = ligand()
    o, ids, sf, gd = oracle_for([lig])
    return o, ids[0], lig, sf, gd


def test_vina_terms_closed_forms(c0):
    o = c0[0]
    # C_H - C_H at the optimal distance 3.8: both gaussians centred, full
    # hydrophobic contact, no repulsion, no hydrogen bond
    assert o.eval_terms(C_H, C_H, 3.8) == pytest.approx(W["g1"] + W["g2"] * np.exp(-2.25) + W["hyd"], abs=1e-15)
    # 1 A inside the optimum: repulsion d^2 = 1, hydrophobic saturated
    e = o.eval_terms(C_H, C_H, 2.8)
    assert e == pytest.approx(W["g1"] * np.exp(-4) + W["g2"] * np.exp(-((-1 - 3) / 2) ** 2) + W["rep"] + W["hyd"],
                              abs=1e-15)
    # donor / acceptor pair (N_D 1.8 + O_A 1.7): h-bond ramps from 0 at d0 to 1 at d0 - 0.7
    d0 = 3.5
    for dx, hb in ((0.0, 0.0), (-0.35, 0.5), (-0.7, 1.0), (-1.0, 1.0)):
        r = d0 + dx
        expect = W["g1"] * np.exp(-(dx / 0.5) ** 2) + W["g2"] * np.exp(-((dx - 3) / 2) ** 2) + W["hb"] * hb
        if dx < 0:
            expect += W["rep"] * dx * dx
        assert o.eval_terms(N_D, O_A, r) == pytest.approx(expect, abs=1e-14)
    # acceptor / acceptor: no h-bond term
    assert o.eval_terms(O_A, O_A, 2.7) == pytest.approx(W["g1"] * np.exp(-(-0.7 / 0.5) ** 2) +
                                                       W["g2"] * np.exp(-((-0.7 - 3) / 2) ** 2) + W["rep"] * 0.49,
                                                       abs=1e-14)


def test_tables_knots_and_midpoints(c0):
    o = c0[0]
    smooth, fast = o.tables()
    n = o.n
    assert n == int(32 * 64) + 3
    for t1, t2 in ((C_H, C_H), (N_D, O_A), (C_P, O_DA)):
        tri = min(t1, t2) + max(t1, t2) * (max(t1, t2) + 1) // 2
        e = smooth[2 * tri * n:2 * (tri + 1) * n:2]
        dor = smooth[2 * tri * n + 1:2 * (tri + 1) * n:2]
        f = fast[tri * n:(tri + 1) * n]
        rs = np.sqrt(np.arange(n) / 32.0)
        for i in (1, 100, 463, 1000, 2048):
            assert e[i] == o.eval_terms(t1, t2, rs[i])  # exact at knots
        assert np.array_equal(f[:-1], (e[1:] + e[:-1]) / 2) and f[-1] == e[-1] / 2
        assert dor[0] == 0 and dor[-1] == 0
        # dor is (de/dr)/r: compare with the analytic derivative away from kinks
        for r in (4.0, 5.5, 7.0):
            i = int(round(r * r * 32))
            h = 1e-6
            de = (o.eval_terms(t1, t2, rs[i] + h) - o.eval_terms(t1, t2, rs[i] - h)) / (2 * h)
            assert dor[i] == pytest.approx(de / rs[i], rel=2e-3, abs=1e-7)


def test_tree_zero_conf_reproduces_input(c0):
    o, m, lig, sf, gd = c0
    e, g, x = o.eval_deriv(m, o.initial_conf(m), (1000, 1000, 1000), backend=1, slope=1e6)
    assert np.max(np.abs(x.reshape(-1, 3) - lig.xyz[:lig.nmovable])) < 1e-12


def quat_mul(l, r):
    a1, b1, c1, d1 = l
    a2, b2, c2, d2 = r
    return np.array([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2])


def rot_quat(v):
    ang = np.linalg.norm(v)
    if ang == 0:
        return np.array([1.0, 0, 0, 0])
    ax = v / ang
    return np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * ax])


def rot_mat(q):
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (a * c + b * d)],
                     [2 * (a * d + b * c), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (a * b + c * d), a * a - b * b - c * c + d * d]])


def test_tree_rigid_motion(c0):
    o, m, lig, sf, gd = c0
    c = o.initial_conf(m)
    c[7:] = [0.3, -1.2, 2.0, 0.1, -0.5, 1.1]
    _, _, x0 = o.eval_deriv(m, c, (1000,) * 3, backend=1)
    x0 = x0.reshape(-1, 3)
    # translating the root moves every atom by the same vector
    c2 = c.copy()
    c2[:3] += [1.5, -2.0, 0.25]
    _, _, x1 = o.eval_deriv(m, c2, (1000,) * 3, backend=1)
    assert np.max(np.abs(x1.reshape(-1, 3) - x0 - [1.5, -2.0, 0.25])) < 1e-12
    # rotating the root rotates every atom about the root origin
    q = rot_quat(np.array([0.4, -0.3, 0.9]))
    c3 = c.copy()
    c3[3:7] = quat_mul(q, c[3:7])
    _, _, x2 = o.eval_deriv(m, c3, (1000,) * 3, backend=1)
    expect = (x0 - c[:3]) @ rot_mat(q).T + c[:3]
    assert np.max(np.abs(x2.reshape(-1, 3) - expect)) < 1e-9
    # internal energy is invariant under rigid motion
    assert o.eval_intramolecular(m, c3) == pytest.approx(o.eval_intramolecular(m, c), abs=1e-9)
    # a full turn of a torsion changes nothing
    c4 = c.copy()
    c4[8] += 2 * np.pi
    _, _, x3 = o.eval_deriv(m, c4, (1000,) * 3, backend=1)
    assert np.max(np.abs(x3 - x0.reshape(-1))) < 1e-9


@pytest.mark.parametrize("backend", [0, 1])
def test_gradient_matches_finite_differences(c0, backend):
    o, m, lig, sf, gd = c0
    f = sg.Flat(lig)
    confs = random_confs(f, gd, 6, seed0=77)
    h = 1e-5
    checked = 0
    for c in confs:
        # keep the ligand compact and inside the box: centre it, moderate torsions
        c = c.copy()
        c[:3] = [0.5, -0.5, 1.0]
        e, g, x = o.eval_deriv(m, c, (1000,) * 3, backend=backend, slope=10)
        if not np.isfinite(e) or e > 1e4:
            continue
        # smina's gradient keeps the receptor term of a box-clamped atom, so it
        # is the derivative of the energy only with every atom inside the box
        x = x.reshape(-1, 3)
        if np.any(x < np.array(gd.begin) + 0.5) or np.any(x > np.array(gd.end) - 0.5):
            continue
        def fd_at(i, h):
            cp, cm = c.copy(), c.copy()
            if i < 3:
                cp[i] += h
                cm[i] -= h
            elif i < 6:
                v = np.zeros(3)
                v[i - 3] = h
                cp[3:7] = quat_mul(rot_quat(v), c[3:7])
                cm[3:7] = quat_mul(rot_quat(-v), c[3:7])
            else:
                cp[i + 1] += h
                cm[i + 1] -= h
            return (o.eval_deriv(m, cp, (1000,) * 3, backend=backend, slope=10)[0] -
                    o.eval_deriv(m, cm, (1000,) * 3, backend=backend, slope=10)[0]) / (2 * h)

        scale = max(1.0, np.max(np.abs(g)))
        for i in range(len(g)):
            a, b = fd_at(i, h), fd_at(i, h / 5)
            # a receptor atom crossing the 8 A cutoff (where Vina's terms are
            # not zero) or a map cell face makes the energy jump: skip those
            if abs(a - b) > 1e-3 * scale:
                continue
            assert abs(b - g[i]) / scale < 2e-2, (i, b, g[i])
            checked += 1
    assert checked >= 3 * 12


@pytest.mark.parametrize("center", [(0.0, 0.0, 0.0), (0.1, 0.2, 0.3)])
def test_grid_equals_brute_force(center):
    """Each map value is the serial sum, in receptor order, of the fast table
    over the heavy receptor atoms within 8 A of the probe and within 8 A of
    the brick of the probe's 3 A cell -- the brick smina's szv_grid_cache
    builds from the first probe of the cell ([floor, ceil] x 3 A of that
    probe).  A box centred on the origin puts those first probes on 3 A
    boundaries (degenerate bricks); the offset box does not, and there the
    value is the plain brute-force sum over every atom within 8 A."""
    lig = ligand()
    gd = sg.grid_dims(center, (22.5, 22.5, 22.5))
    o, _, _, _ = oracle_for([lig], gd=gd)
    rec = receptor()
    smooth, fast = o.tables()
    n = o.n
    heavy = rec.type > 1
    rx, rt = rec.xyz[heavy], rec.type[heavy]
    t = C_H
    grid = o.grid(t).reshape(gd.n[2] + 1, gd.n[1] + 1, gd.n[0] + 1)
    fi = [(gd.end[i] - gd.begin[i]) / gd.n[i] for i in range(3)]
    rng = np.random.default_rng(3)
    exact = 0
    for _ in range(300):
        ix = rng.integers(0, gd.n[0] + 1, 3)
        # smina's probe coordinates: init + (1 / factor) * index
        p = np.array([gd.begin[i] + (1 / (gd.n[i] / (gd.end[i] - gd.begin[i]))) * ix[i] for i in range(3)])
        # first probe of the point's 3 A cell, per dimension: its brick
        lo, hi, degenerate = [], [], False
        for d in range(3):
            cell = np.floor(p[d] / 3.0)
            j = 0
            while np.floor((gd.begin[d] + (1 / (gd.n[d] / (gd.end[d] - gd.begin[d]))) * j) / 3.0) != cell:
                j += 1
            first = gd.begin[d] + (1 / (gd.n[d] / (gd.end[d] - gd.begin[d]))) * j
            lo.append(np.floor(first / 3) * 3)
            hi.append(np.ceil(first / 3) * 3)
            degenerate |= lo[-1] == hi[-1]
        s = 0.0
        for a in range(len(rt)):
            d2 = (rx[a, 0] - p[0]) ** 2 + (rx[a, 1] - p[1]) ** 2 + (rx[a, 2] - p[2]) ** 2
            if d2 <= 64:
                closest = np.clip(rx[a], lo, hi)
                if np.sum((closest - rx[a]) ** 2) < 64:  # in the cell's atom list
                    tri = min(rt[a], t) + max(rt[a], t) * (max(rt[a], t) + 1) // 2
                    s += fast[tri * n + int(32 * d2)]
        assert grid[ix[2], ix[1], ix[0]] == s
        if not degenerate:
            full = 0.0
            for a in range(len(rt)):
                d2 = (rx[a, 0] - p[0]) ** 2 + (rx[a, 1] - p[1]) ** 2 + (rx[a, 2] - p[2]) ** 2
                if d2 <= 64:
                    tri = min(rt[a], t) + max(rt[a], t) * (max(rt[a], t) + 1) // 2
                    full += fast[tri * n + int(32 * d2)]
            assert grid[ix[2], ix[1], ix[0]] == full
            exact += 1
    if center == (0.1, 0.2, 0.3):
        assert exact > 250


def test_bfgs_reaches_quadratic_minimum():
    rng = np.random.default_rng(0)
    for n in (3, 8, 20):
        M = rng.normal(size=(n, n))
        A = M @ M.T + n * np.eye(n)
        b = rng.normal(size=n) * 3
        x_star = np.linalg.solve(A, b)
        f_star = -0.5 * b @ x_star
        for accurate in (False, True):
            f, x, ev = po.bfgs_quadratic(A, b, np.zeros(n), 200, accurate)
            g = A @ x - b
            assert g @ g < 1e-4  # smina's stopping rule
            assert f >= f_star - 1e-12 and f - f_star < 1e-4
            assert ev < 200


def test_mc_chain_is_deterministic_and_sorted(c0):
    o, m, lig, sf, gd = c0
    f = sg.Flat(lig)
    c, s = f.make_chain(4242, gd)
    p = sg.mc_params()
    e1, c1, x1, ev1 = o.mc_chain(m, p, 30, (25 + f.natoms) // 3, c, s)
    e2, c2, x2, ev2 = o.mc_chain(m, p, 30, (25 + f.natoms) // 3, c, s)
    assert np.array_equal(e1, e2) and np.array_equal(c1, c2) and ev1 == ev2
    assert np.all(np.diff(e1) >= 0) and 1 <= len(e1) <= 20
    # a one-step chain keeps exactly its one accepted, refined pose
    e3, _, _, ev3 = o.mc_chain(m, p, 1, (25 + f.natoms) // 3, c, s)
    assert len(e3) == 1 and ev3 > 2
    e0, _, _, ev0 = o.mc_chain(m, p, 0, (25 + f.natoms) // 3, c, s)
    assert len(e0) == 0 and ev0 == 0


def spline_at(coef, cut, tri, r, fraction):
    """Spline::eval_deriv (src/lib/splines.h:103-121) in numpy"""
    if r >= cut[tri]:
        return 0.0, 0.0
    x, a, b, c, d = coef[tri, int(r / fraction)]
    lx = r - x
    return ((a * lx + b) * lx + c) * lx + d, (3 * a * lx + 2 * b) * lx + c


def test_splines_interpolate_the_terms():
    """precalculate_splines (src/lib/precalculate.h:313-455, factor 10): the
    cubic spline through the terms at r = i/10 plus (8, 0) reproduces them
    at the knots, is continuous with its first and second derivatives
    there, has zero slope at both ends, and between the knots of a smooth
    stretch (C_H pair beyond the hydrophobic ramp: gaussians only) stays
    within 1e-6 of the terms and of their derivative.  (The kink at the end
    of the hydrophobic ramp, r = 5.3, and the pinned zero at the cutoff bend
    the spline away from the terms near them, by design.)"""
    lig = ligand()
    o, ids, sf, gd = oracle_for([lig])
    coef, cut = o.use_splines(10.0)
    h = 0.1
    for t1, t2 in ((C_H, C_H), (N_D, O_A), (C_P, O_DA)):
        tri = min(t1, t2) + max(t1, t2) * (max(t1, t2) + 1) // 2
        assert cut[tri] == 8.0
        s = coef[tri]
        assert np.array_equal(s[:, 0], np.arange(80) * h)
        for i in range(80):
            assert s[i, 4] == o.eval_terms(t1, t2, i * h)  # the knots themselves
        # the end of each segment meets the next one: value, slope, curvature
        a, b, c, d = s[:-1, 1], s[:-1, 2], s[:-1, 3], s[:-1, 4]
        assert np.allclose(((a * h + b) * h + c) * h + d, s[1:, 4], rtol=0, atol=1e-9 * max(1, np.abs(d).max()))
        assert np.allclose((3 * a * h + 2 * b) * h + c, s[1:, 3], rtol=0, atol=1e-8 * max(1, np.abs(c).max()))
        assert np.allclose(6 * a * h + 2 * b, 2 * s[1:, 2], rtol=0, atol=1e-7 * max(1, np.abs(b).max()))
        # clamped ends: zero slope at 0 and at the cutoff, value 0 there
        assert abs(s[0, 3]) < 1e-9 * max(1, np.abs(s[:, 3]).max())
        x, a, b, c, d = s[-1]
        assert abs(((a * h + b) * h + c) * h + d) < 1e-9 and abs((3 * a * h + 2 * b) * h + c) < 1e-8
    tri = C_H + C_H * (C_H + 1) // 2
    for r in np.arange(6.05, 7.0, 0.1):
        v, dv = spline_at(coef, cut, tri, r, h)
        exact = o.eval_terms(C_H, C_H, r)
        slope = (o.eval_terms(C_H, C_H, r + 1e-6) - o.eval_terms(C_H, C_H, r - 1e-6)) / 2e-6
        assert abs(v - exact) < 1e-6 and abs(dv - slope) < 1e-6, r


def test_spline_eval_deriv_matches_the_tables():
    """the oracle's direct-receptor evaluation with splines: the energy
    moves off the linear tables' (the splines reach zero at the cutoff,
    where the terms do not), and the gradient is the finite difference of
    the energy (the splines are C2)"""
    lig = ligand()
    o, ids, sf, gd = oracle_for([lig])
    # the parsed pose, moved a little: every atom inside the box, where the
    # direct term is smooth (outside it, smina adds the clamped point's
    # derivative to the penalty's)
    conf = o.initial_conf(ids[0])
    conf[:3] += (0.3, -0.2, 0.1)
    e_lin, g_lin, _ = o.eval_deriv(ids[0], conf, (1000.0, 1000.0, 1000.0), backend=1)
    o.use_splines(10.0)
    e_spl, g_spl, _ = o.eval_deriv(ids[0], conf, (1000.0, 1000.0, 1000.0), backend=1)
    assert e_spl != e_lin and abs(e_spl - e_lin) < 1.0
    # finite differences along the position coordinates
    for k in range(3):
        cp, cm = conf.copy(), conf.copy()
        cp[k] += 1e-6
        cm[k] -= 1e-6
        fd = (o.eval_deriv(ids[0], cp, (1000.0, 1000.0, 1000.0), backend=1)[0] -
              o.eval_deriv(ids[0], cm, (1000.0, 1000.0, 1000.0), backend=1)[0]) / 2e-6
        assert abs(fd - g_spl[k]) < 1e-4 * max(1.0, abs(fd)), k

"""Known-answer tests of the oracle's Monte Carlo step: mutation, the
Metropolis rule and the output container, predicted from smina's rules
without the oracle's search code.

The chains run with maxiters 0, so BFGS changes no conf and every candidate
is the mutated conf itself (src/lib/bfgs.h:203-281: no iteration, f0 kept).
The random stream is restated here in Python (MT19937 from its published
recurrence, cross-checked against numpy's MT19937, and Boost's uniform_real /
uniform_int rules), and each step's outcome is predicted from it:

  mutate_conf (src/lib/mutate.cpp:35-63): which = random_int(0, 2 + T - 1);
    0: position += 2 x random_inside_sphere (src/lib/random.cpp:66-76);
    1: orientation <- q(2 / r_g x random_inside_sphere) * orientation, r_g the
       heavy atoms' gyration radius about the root origin in the model's
       current coordinates (src/lib/model.cpp:868-883);
    2..: torsion <- random_fl(-pi, pi);
  metropolis_accept (src/lib/monte_carlo.cpp:37-41): new < old, else one
    draw u = random_fl(0, 1) and u < exp((old - new) / 1.2);
  add_to_output_container (src/lib/coords.cpp:34-58): replace the closest
    pose if within 1 A and better, else append (20 at most: replace the
    worst if better), sorted by energy.

Only energies and coordinates come from the oracle's evaluation (pinned by
the finite-difference and closed-form tests elsewhere)."""
import math

import numpy as np
import pytest

import pyoracle as po
import smina_gpu as sg
from common import gd_c0, ligand, oracle_for

TEMPERATURE, AMPLITUDE, HUNT, AUTHENTIC = 1.2, 2.0, 10.0, 1000.0


class MT:
    """MT19937 from its published recurrence (624 words, twist, temper)"""

    def __init__(self, state):
        self.x = [int(w) for w in state[:624]]
        self.i = int(state[624])

    def next(self):
        if self.i >= 624:
            x = self.x
            for k in range(624):
                y = (x[k] & 0x80000000) | (x[(k + 1) % 624] & 0x7fffffff)
                x[k] = x[(k + 397) % 624] ^ (y >> 1) ^ (0x9908b0df if y & 1 else 0)
            self.i = 0
        z = self.x[self.i]
        self.i += 1
        z ^= z >> 11
        z ^= (z << 7) & 0x9d2c5680
        z ^= (z << 15) & 0xefc60000
        z ^= z >> 18
        return z & 0xffffffff

    def fl(self, a, b):
        """boost::uniform_real<double>(a, b): w / 2^32 * (b - a) + a, again if >= b"""
        while True:
            r = self.next() / 4294967296.0 * (b - a) + a
            if r < b:
                return r

    def int(self, a, b):
        """boost::uniform_int<int>(a, b): bucket rejection over 2^32 - 1"""
        rng = (b - a) & 0xffffffff
        if rng == 0:
            return a
        bucket = 0xffffffff // (rng + 1)
        if 0xffffffff % (rng + 1) == rng:
            bucket += 1
        while True:
            v = self.next() // bucket
            if v <= rng:
                return v + a

    def sphere(self):
        while True:
            v = np.array([self.fl(-1.0, 1.0) for _ in range(3)])
            if float(v @ v) < 1:
                return v


def test_python_mt_matches_numpy():
    """the restated generator against numpy's MT19937 from the same state"""
    state = np.zeros(625, np.uint32)
    state[0] = 5489
    for k in range(1, 624):
        p = int(state[k - 1])
        state[k] = (1812433253 * (p ^ (p >> 30)) + k) & 0xffffffff
    state[624] = 624
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": state[:624].copy(), "pos": 624}}
    ref = bg.random_raw(2000)
    mt = MT(state)
    assert [mt.next() for _ in range(2000)] == [int(v) for v in ref]
    # the C++ standard's known answer for the default seed
    mt = MT(state)
    for _ in range(9999):
        mt.next()
    assert mt.next() == 4123659995


def qmul(a, b):
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                     a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                     a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])


def rotation_quaternion(rot):
    ang = float(np.linalg.norm(rot))
    axis = rot / ang
    while ang > math.pi:
        ang -= 2 * math.pi
    return np.array([math.cos(ang / 2), *(math.sin(ang / 2) * axis)])


@pytest.fixture(scope="module")
def chain_setup():
    lig = ligand(1, 24, 6)
    o, (m,), sf, gd = oracle_for([lig])
    f = sg.Flat(lig)
    heavy = np.array([t > 1 for t in lig.type[:lig.nmovable]])
    return o, m, f, gd, lig, heavy


def mutate_prediction(mt, conf, ntors, gyration):
    """the mutated conf and the branch taken"""
    c = conf.copy()
    which = mt.int(0, 2 + ntors - 1)
    if which == 0:
        c[0:3] = c[0:3] + AMPLITUDE * mt.sphere()
        return c, "position"
    if which == 1:
        rot = (AMPLITUDE / gyration) * mt.sphere()
        q = qmul(rotation_quaternion(rot), c[3:7])
        c[3:7] = q / np.linalg.norm(q)
        return c, "orientation"
    c[7 + which - 2] = mt.fl(-math.pi, math.pi)
    return c, f"torsion {which - 2}"


def gyration_radius(xyz, heavy, origin):
    d = xyz[heavy] - origin
    return math.sqrt(float(np.mean(np.sum(d * d, axis=1))))


def assert_conf(got, want, branch):
    if branch == "orientation":
        assert np.array_equal(got[:3], want[:3]) and np.array_equal(got[7:], want[7:])
        assert np.allclose(got[3:7], want[3:7], rtol=0, atol=1e-12), (got[3:7], want[3:7])
    else:
        assert np.array_equal(got, want), (branch, got - want)


def test_one_step_saves_the_mutated_conf(chain_setup):
    """steps 1, maxiters 0: the single saved pose is the mutation of the
    start conf, with its authentic-cap energy; every branch is reached"""
    o, m, f, gd, lig, heavy = chain_setup
    params = sg.mc_params()
    origin0 = f.initial_conf()[0:3]
    # the model's coordinates before the first mutation are the parsed ones
    gyr0 = gyration_radius(lig.xyz[:lig.nmovable], heavy, origin0)
    seen = set()
    for seed in range(7000, 7060):
        c0, s0 = f.make_chain(seed, gd)
        mt = MT(s0)
        want, branch = mutate_prediction(mt, c0, f.ntors_total, gyr0)
        e, confs, _, _ = o.mc_chain(m, params, 1, 0, c0, s0)
        assert len(e) == 1
        assert_conf(confs[0], want, branch)
        e_auth = o.eval_deriv(m, confs[0], (AUTHENTIC,) * 3)[0]
        assert e[0] == pytest.approx(e_auth, rel=1e-12, abs=1e-12)
        seen.add(branch.split()[0])
    assert seen == {"position", "orientation", "torsion"}


def test_two_steps_follow_metropolis_and_the_container(chain_setup):
    """steps 2, maxiters 0: the second mutation starts from the first; it is
    accepted when its hunt-cap energy is lower than the current authentic-cap
    energy or when u < exp((old - new) / 1.2); an accepted pose enters the
    container by smina's rule.  The seeds reach acceptance both ways,
    rejection, appending and replacement within 1 A."""
    o, m, f, gd, lig, heavy = chain_setup
    params = sg.mc_params()
    origin0 = f.initial_conf()[0:3]
    gyr0 = gyration_radius(lig.xyz[:lig.nmovable], heavy, origin0)
    # start confs near minima (a short search's saved poses), where a
    # mutation changes the energy by a few kcal/mol and the Metropolis draw
    # decides; the random stream is each seed's chain stream
    starts = []
    for k in range(4):
        c, st = f.make_chain(6000 + k, gd)
        starts += list(o.mc_chain(m, params, 40, 12, c, st)[1][:5])
    outcomes = set()
    # draws whose outcome would differ at a temperature of 1.1 or 1.3
    sensitive = {1.1: 0, 1.3: 0}
    for seed in range(7100, 9100):
        _, s0 = f.make_chain(seed, gd)
        c0 = starts[seed % len(starts)]
        mt = MT(s0)
        c1, b1 = mutate_prediction(mt, c0, f.ntors_total, gyr0)
        e1, _, x1 = o.eval_deriv(m, c1, (AUTHENTIC,) * 3)
        # the model now holds c1's coordinates (set after the step is taken)
        x1 = x1.reshape(-1, 3)
        mheavy = np.array(heavy_model_mask(o, m, f))
        gyr1 = gyration_radius(x1, mheavy, c1[0:3])
        c2, b2 = mutate_prediction(mt, c1, f.ntors_total, gyr1)
        e2_hunt = o.eval_deriv(m, c2, (HUNT,) * 3)[0]
        if e2_hunt < e1:
            accepted, how = True, "lower"
        else:
            u = mt.fl(0.0, 1.0)
            accepted = u < math.exp((e1 - e2_hunt) / TEMPERATURE)
            how = "draw" if accepted else "rejected"
            for t in sensitive:
                sensitive[t] += (u < math.exp((e1 - e2_hunt) / t)) != accepted
        e, confs, xs, _ = o.mc_chain(m, params, 2, 0, c0, s0)
        if not accepted:
            assert len(e) == 1 and confs[0] == pytest.approx(c1, abs=1e-12)
            outcomes.add("rejected")
            continue
        e2, _, x2 = o.eval_deriv(m, c2, (AUTHENTIC,) * 3)
        x2 = x2.reshape(-1, 3)[mheavy]
        rmsd = math.sqrt(float(np.mean(np.sum((x2 - x1[mheavy]) ** 2, axis=1))))
        if rmsd < 1.0:
            want = [(e2, c2)] if e2 < e1 else [(e1, c1)]
            kind = "replaced" if e2 < e1 else "kept"
        else:
            want = sorted([(e1, c1), (e2, c2)], key=lambda t: t[0])
            kind = "appended"
        assert len(e) == len(want), (seed, how, kind)
        for k, (ew, cw) in enumerate(want):
            assert e[k] == pytest.approx(ew, rel=1e-12, abs=1e-12)
            assert confs[k] == pytest.approx(cw, abs=1e-12)
        outcomes.add(how)
        outcomes.add(kind)
    assert {"lower", "draw", "rejected", "appended"} <= outcomes, outcomes
    assert outcomes & {"replaced", "kept"}, outcomes
    assert min(sensitive.values()) >= 1, sensitive


def heavy_model_mask(o, m, f):
    """heavy atoms in the model's atom order (the oracle's coordinates)"""
    t = np.ctypeslib.as_array(f.view.atom_type, (f.natoms,))
    return [int(v) > 1 for v in t]


def _pose(tag, e, x):
    return (tag, e, np.asarray(x, dtype=np.float64))


def test_container_rules_by_hand():
    """hand-built pose sets of one atom: append, replace within 1 A when
    better, keep when worse, replace the worst of a full container, discard
    a worse pose when full, first index on equal distances, sorted output"""
    at = lambda x: [[x, 0.0, 0.0]]
    out = po.container_add([], _pose(0, -5.0, at(0.0)), max_size=3)
    assert [t for t, _, _ in out] == [0]
    out = po.container_add(out, _pose(1, -4.0, at(5.0)), max_size=3)       # far: appended, sorted
    assert [t for t, _, _ in out] == [0, 1]
    out = po.container_add(out, _pose(2, -6.0, at(5.5)), max_size=3)       # 0.5 A from 1, better: replaces it
    assert [t for t, _, _ in out] == [2, 0]
    out = po.container_add(out, _pose(3, -1.0, at(0.9)), max_size=3)       # 0.9 A from 0, worse: discarded
    assert [t for t, _, _ in out] == [2, 0]
    out = po.container_add(out, _pose(4, -2.0, at(-9.0)), max_size=3)      # far, room left: appended
    assert [t for t, _, _ in out] == [2, 0, 4]
    out = po.container_add(out, _pose(5, -3.0, at(20.0)), max_size=3)      # full, far, better than the worst
    assert [t for t, _, _ in out] == [2, 0, 5]
    out = po.container_add(out, _pose(6, -0.5, at(-20.0)), max_size=3)     # full, far, worse than the worst
    assert [t for t, _, _ in out] == [2, 0, 5]
    # equal distance (0.5 A) to two poses: the one first in energy order is replaced
    two = [_pose(10, -3.0, at(0.0)), _pose(11, -2.0, at(1.0))]
    out = po.container_add(two, _pose(12, -9.0, at(0.5)), max_size=5)
    assert [t for t, _, _ in out] == [12, 11]
    # exactly 1 A is not within min_rmsd: appended
    out = po.container_add([_pose(20, -3.0, at(0.0))], _pose(21, -9.0, at(1.0)), max_size=5)
    assert [t for t, _, _ in out] == [21, 20]

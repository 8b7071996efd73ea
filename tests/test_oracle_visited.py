"""Pins of the oracle's QuickVina 2 history check (CPU), the check the
GPU-style --fast search runs before each local optimization
(src/gpu/gpu_mc.cuh:146-264; QuickVina 2, Alhossary et al., Bioinformatics
2015).  Every expected answer below is worked out by hand from the
published rule, not from the oracle:

- until the ring of 5n points is full, every point is interesting;
- otherwise the 4n stored points nearest to x, in order of (squared
  distance, slot), are checked, and the first that passes makes x
  interesting;
- a stored point passes unless some coordinate in which both gradients have
  the same nonzero sign breaks (g > 0) == (x > x_e) xor (f > f_e).

With n = 1, x = 0, g = +1, f = 0 and stored gradients +1, a point at
x_e > 0 passes iff f_e < 0, and a point at x_e < 0 passes iff f_e >= 0."""
import numpy as np

import pyoracle as po


def check(points, energies, grads, x=(0.0,), f=0.0, g=(1.0,)):
    n = len(x)
    return po.visited_check(np.reshape(points, (-1, n)), energies, np.reshape(grads, (-1, n)), x, f, g)


def test_not_full_is_always_interesting():
    # 4 of 5 slots, none passing
    assert check([1, 2, 3, 4], [1, 1, 1, 1], [1, 1, 1, 1])


def test_full_and_nothing_passes():
    assert not check([1, 2, 3, 4, 5], [1] * 5, [1] * 5)


def test_passing_point_within_the_4n_nearest():
    # only x_e = 4 (rank 3 of 0..4) passes
    assert check([1, 2, 3, 4, 5], [1, 1, 1, -1, 1], [1] * 5)
    # the slot order does not matter, only the distance
    assert check([4, 5, 1, 3, 2], [-1, 1, 1, 1, 1], [1] * 5)


def test_passing_point_beyond_the_4n_nearest():
    # only x_e = 5 (rank 4) passes: it is never checked
    assert not check([1, 2, 3, 4, 5], [1, 1, 1, 1, -1], [1] * 5)


def test_opposite_gradient_signs_skip_the_coordinate():
    # x_e = 1 would fail (f_e > 0), but its gradient sign differs: vacuous pass
    assert check([1, 2, 3, 4, 5], [1] * 5, [-1, 1, 1, 1, 1])
    # a zero stored gradient skips the coordinate too
    assert check([1, 2, 3, 4, 5], [1] * 5, [0, 1, 1, 1, 1])


def test_zero_gradient_now_makes_every_point_pass():
    assert check([1, 2, 3, 4, 5], [1] * 5, [1] * 5, g=(0.0,))


def test_points_on_the_other_side():
    # x_e < 0 passes iff f_e >= 0
    assert check([-1, 2, 3, 4, 5], [1, 1, 1, 1, 1], [1] * 5)
    assert not check([-1, 2, 3, 4, 5], [-1, 1, 1, 1, 1], [1] * 5)


def test_distance_ties_go_to_the_lower_slot():
    # x_e = -4 and +4 tie at rank 3/4: only -4 passes (f_e >= 0)
    assert check([1, 2, 3, -4, 4], [1] * 5, [1] * 5)
    assert not check([1, 2, 3, 4, -4], [1] * 5, [1] * 5)


def test_ring_overwrites_the_oldest_point():
    # x_e = 1 passes; a sixth point takes its slot
    assert check([1, 2, 3, 4, 5], [-1, 1, 1, 1, 1], [1] * 5)
    assert not check([1, 2, 3, 4, 5, 6], [-1, 1, 1, 1, 1, 1], [1] * 6)
    # after a full turn the ring again holds the last 5n points
    assert check([9, 9, 9, 9, 9, 1, 2, 3, 4, 5], [1] * 5 + [-1, 1, 1, 1, 1], [1] * 10)


def test_every_coordinate_must_be_consistent():
    # n = 2, x = (0, 0), g = (+1, +1), f = 0; 9 far failing points and one
    # near point with f_e = -1: at (1, 1) both coordinates agree, at (1, -1)
    # the second breaks the rule
    far = [[10.0 + k, 10.0 + k] for k in range(9)]
    grads = [[1.0, 1.0]] * 10
    energies = [-1.0] + [1.0] * 9
    assert check([[1.0, 1.0]] + far, energies, grads, x=(0.0, 0.0), g=(1.0, 1.0))
    assert not check([[1.0, -1.0]] + far, energies, grads, x=(0.0, 0.0), g=(1.0, 1.0))

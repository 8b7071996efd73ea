"""Pins of the oracle's arithmetic building blocks: the portable sin / cos /
exp / acos (fdlibm method) against glibc through numpy within 1 ulp, exact
special values; MT19937 against numpy's independent implementation and the
C++ standard's known answer; the uniform distributions against their
range and rejection invariants."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pm():
    """portable math and the product's MT19937, compiled from the shared
    headers into a tiny test library"""
    src = os.path.join(ROOT, "tests", "native", "pm_probe.cpp")
    out = os.path.join(ROOT, "tests", "native", "_pm_probe.so")
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        assert os.system(f"g++ -O2 -std=c++17 -ffp-contract=off -fPIC -shared -o {out} {src}") == 0
    lib = C.CDLL(out)
    d = C.POINTER(C.c_double)
    for name in ("pm_sin", "pm_cos", "pm_exp", "pm_acos"):
        getattr(lib, name).argtypes = [d, d, C.c_int]
    lib.mt_stream.argtypes = [C.c_uint32, C.POINTER(C.c_uint32), C.c_int]
    lib.uniform_fl.argtypes = [C.c_uint32, C.c_double, C.c_double, d, C.c_int]
    lib.uniform_int.argtypes = [C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
    return lib


def ulps(a, b):
    ia = a.view(np.int64).astype(np.int64)
    ib = b.view(np.int64).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-(2 ** 63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2 ** 63)) - ib, ib)
    return np.abs(ia - ib)


def call(lib, name, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    getattr(lib, name)(x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)), len(x))
    return y


@pytest.mark.parametrize("name,ref,lo,hi", [
    ("pm_sin", np.sin, -20, 20), ("pm_cos", np.cos, -20, 20), ("pm_exp", np.exp, -745, 709),
    ("pm_acos", np.arccos, -1, 1)])
def test_portable_math_within_one_ulp(pm, name, ref, lo, hi):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(lo, hi, 400000), np.linspace(lo, hi, 100001)])
    if name in ("pm_sin", "pm_cos"):  # the search's own range, densely
        x = np.concatenate([x, rng.uniform(-np.pi / 2, np.pi / 2, 200000)])
    assert ulps(call(pm, name, x), ref(x)).max() <= 1


def test_portable_math_special_values(pm):
    assert call(pm, "pm_exp", [0.0])[0] == 1.0
    assert call(pm, "pm_cos", [0.0])[0] == 1.0
    assert call(pm, "pm_sin", [0.0])[0] == 0.0
    assert call(pm, "pm_acos", [1.0])[0] == 0.0
    assert call(pm, "pm_acos", [-1.0])[0] == np.pi
    assert call(pm, "pm_exp", [-800.0])[0] == 0.0
    assert np.isinf(call(pm, "pm_exp", [710.0])[0])
    # odd / even symmetry
    x = np.linspace(0.01, 3.0, 1000)
    assert np.array_equal(call(pm, "pm_sin", -x), -call(pm, "pm_sin", x))
    assert np.array_equal(call(pm, "pm_cos", -x), call(pm, "pm_cos", x))


def test_mt19937_against_numpy_and_known_answer(pm):
    n = 10000
    out = np.zeros(n, np.uint32)
    pm.mt_stream(5489, out.ctypes.data_as(C.POINTER(C.c_uint32)), n)
    # C++ [rand.predef]: the 10000th output of default-seeded mt19937
    assert int(out[-1]) == 4123659995
    for seed in (1, 12345, 2 ** 31 + 7):
        bg = np.random.MT19937()
        bg._legacy_seeding(seed)
        ref = bg.random_raw(3000).astype(np.uint32)
        got = np.zeros(3000, np.uint32)
        pm.mt_stream(seed, got.ctypes.data_as(C.POINTER(C.c_uint32)), 3000)
        assert np.array_equal(got, ref)


def test_uniform_distribution_invariants(pm):
    # random_fl(a, b) = u / 2^32 * (b - a) + a, rejected when >= b
    n = 20000
    y = np.zeros(n)
    pm.uniform_fl(3, -np.pi, np.pi, y.ctypes.data_as(C.POINTER(C.c_double)), n)
    assert y.min() >= -np.pi and y.max() < np.pi
    raw = np.zeros(n, np.uint32)
    pm.mt_stream(3, raw.ctypes.data_as(C.POINTER(C.c_uint32)), n)
    expect = raw / 4294967296.0 * (2 * np.pi) + -np.pi
    acc = expect[expect < np.pi]
    m = min(len(acc), n)
    assert m > n - 10 and np.array_equal(y[:m], acc[:m])
    # random_int(a, b): bucket rejection, every value reachable, none outside
    k = np.zeros(n, np.int32)
    pm.uniform_int(9, 0, 6, k.ctypes.data_as(C.POINTER(C.c_int)), n)
    assert k.min() == 0 and k.max() == 6
    counts = np.bincount(k, minlength=7)
    assert np.all(np.abs(counts / n - 1 / 7) < 0.02)
    # range 0..10^6 (parallel_mc's task seeds)
    pm.uniform_int(1, 0, 1000000, k.ctypes.data_as(C.POINTER(C.c_int)), 1000)
    assert k[:1000].min() >= 0 and k[:1000].max() <= 1000000
    bucket = 0xFFFFFFFF // 1000001
    raw = np.zeros(1010, np.uint32)
    pm.mt_stream(1, raw.ctypes.data_as(C.POINTER(C.c_uint32)), 1010)
    acc = [int(r) // bucket for r in raw if int(r) // bucket <= 1000000][:1000]
    assert list(k[:1000]) == acc

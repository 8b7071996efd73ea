"""The build's portable math and random stream against the reference's own
headers, bit for bit.

oracle/Makefile's `ref` target compiles src/lib/portable_math.h (the sin,
cos, exp and acos of smina's SMINA_PORTABLE_MATH build and of its FP64 GPU
kernels) and src/gpu/gpu_rng.cuh (its device restatement of boost::mt19937
and Boost's uniform_real / uniform_int) from the reference tree, beside
gpu-smina_amd/csrc/pmath.h and rng.h, into oracle/_ref/.  The comparisons
run in C++ (oracle/ref_pin.cpp); every result must have the same bits.

CPU only.  Skipped where the reference tree is absent and nothing was built.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_pin.so")
REFERENCE = "/root/reference"
N = 10_000_000  # random arguments per function and range


@pytest.fixture(scope="module")
def pin():
    if os.path.isdir(REFERENCE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if not os.path.exists(LIB):
        pytest.skip("the reference tree is absent and oracle/_ref was not built")
    lib = C.CDLL(LIB)
    lib.pin_math.restype = C.c_longlong
    lib.pin_math.argtypes = [C.c_int, C.c_double, C.c_double, C.c_longlong, C.c_uint64, C.POINTER(C.c_double)]
    lib.pin_math_at.restype = C.c_longlong
    lib.pin_math_at.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_longlong, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    lib.pin_rng.restype = C.c_longlong
    lib.pin_rng.argtypes = [C.c_int, C.c_uint32, C.c_longlong, C.c_double, C.c_double, C.c_int, C.c_int]
    return lib


FNS = {"sin": 0, "cos": 1, "exp": 2, "acos": 3}


@pytest.mark.parametrize("fn,lo,hi", [
    ("sin", -math.pi, math.pi),          # the search's torsion and half-angle range
    ("cos", -math.pi, math.pi),
    ("sin", -4 * math.pi, 4 * math.pi),  # past one reduction quadrant cycle
    ("cos", -4 * math.pi, 4 * math.pi),
    ("sin", math.pi / 4, math.pi / 2),   # reduced to the cos kernel near its qx split
    ("cos", 0.25, 0.8),                  # the cos kernel's three pieces: < 0.3, qx, > 0.78125
    ("exp", -50.0, 0.0),                 # Metropolis: exp((old - new) / 1.2), old < new
    ("exp", -745.5, 710.0),              # the whole finite range, overflow and underflow
    ("exp", -0.7, 0.7),                  # around the ln2/2 reduction threshold
    ("acos", -1.0, 1.0),
])
def test_portable_math_bits_equal_reference(pin, fn, lo, hi):
    first = C.c_double(float("nan"))
    bad = pin.pin_math(FNS[fn], lo, hi, N, 12345 + FNS[fn], C.byref(first))
    assert bad == 0, f"{fn} differs from portable_math on {bad} of {N} arguments in [{lo}, {hi}), first {first.value!r}"


def _edges():
    """exact thresholds of the reference's branches, each with its
    neighbouring doubles"""
    pts = [0.0, -0.0, 0.3, 0.78125, 0.7853981633974483, math.pi / 2, math.pi, 2 * math.pi, 1e-300, 5e-324,
           0.5 * 0.6931471805599453, 3.7252902984e-09, 2.0 ** -28, 709.782712893384, -745.1332191019411,
           0.5, -0.5, 1.0, -1.0, 1e-17, -1e-17]
    out = []
    for p in pts:
        for s in (1.0, -1.0):
            v = s * p
            out += [v, np.nextafter(v, np.inf), np.nextafter(v, -np.inf)]
    for k in range(-16, 17):
        v = k * math.pi / 4
        out += [v, np.nextafter(v, np.inf), np.nextafter(v, -np.inf)]
    return np.array(out, dtype=np.float64)


@pytest.mark.parametrize("fn", list(FNS))
def test_portable_math_edges_equal_reference(pin, fn):
    x = _edges()
    if fn == "acos":
        x = x[np.abs(x) <= 1.0]
    a = np.zeros_like(x)
    b = np.zeros_like(x)
    dp = C.POINTER(C.c_double)
    bad = pin.pin_math_at(FNS[fn], x.ctypes.data_as(dp), len(x), a.ctypes.data_as(dp), b.ctypes.data_as(dp))
    diff = [(float(x[i]), float(a[i]), float(b[i])) for i in range(len(x))
            if a[i].tobytes() != b[i].tobytes() and not (np.isnan(a[i]) and np.isnan(b[i]))]
    assert bad == 0 and not diff, diff[:5]


@pytest.mark.parametrize("kind,args", [
    (0, (0.0, 0.0, 0, 0)),                     # raw tempered words across many twists
    (1, (0.0, 1.0, 0, 0)),                     # Metropolis' uniform(0, 1)
    (1, (-math.pi, math.pi, 0, 0)),            # torsion mutation
    (1, (-1.0, 1.0, 0, 0)),                    # random_inside_sphere
    (2, (0.0, 0.0, 0, 999999)),                # task seeds random_int(0, 1e6)
    (2, (0.0, 0.0, 0, 8)),                     # which DOF to mutate
    (2, (0.0, 0.0, -3, 0x7ffffffe)),           # a range past 2^31
])
def test_random_stream_equals_reference(pin, kind, args):
    for seed in (0, 1, 5489, 123456789, 4294967295):
        bad = pin.pin_rng(kind, seed, 200_000, *args)
        assert bad == 0, (kind, seed, bad)

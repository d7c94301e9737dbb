"""The 64.64 fixed-point words behind the GPU's exact force sums
(melt_ibm_amd/csrc/fixed128.h), compiled for the host and checked against
exact rational arithmetic: the split must equal trunc(v * 2^64) for every
double (tiny negatives and values next to integers included), sums must not
depend on order, and the conversion back must be within one ulp."""
import ctypes as C
import math
import random
import shutil
import subprocess
from fractions import Fraction
from pathlib import Path

import pytest

HDR = Path(__file__).resolve().parents[1] / "melt_ibm_amd" / "csrc" / "fixed128.h"

SRC = r"""
#include "fixed128.h"
extern "C" void split(double v, unsigned long long* lo, unsigned long long* hi) { acc_split(v, *lo, *hi); }
extern "C" double value(unsigned long long lo, unsigned long long hi) { return acc_value(lo, hi); }
extern "C" void sum(const double* v, int n, unsigned long long* lo, unsigned long long* hi) {
  unsigned long long a = 0, b = 0;
  for (int i = 0; i < n; ++i) { unsigned long long l, h; acc_split(v[i], l, h); acc_add_words(a, b, l, h); }
  *lo = a; *hi = b;
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("fx")
    (d / "fx.cpp").write_text(SRC)
    so = d / "libfx.so"
    subprocess.run([cxx, "-O2", "-shared", "-fPIC", "-DMELT_HD=", f"-I{HDR.parent}", str(d / "fx.cpp"), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.split.argtypes = [C.c_double, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    L.value.argtypes = [C.c_ulonglong, C.c_ulonglong]
    L.value.restype = C.c_double
    L.sum.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    return L


def words(L, v):
    lo, hi = C.c_ulonglong(), C.c_ulonglong()
    L.split(v, C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


def exact(v):
    return int(Fraction(v) * 2**64) % 2**128  # int() truncates toward zero


def signed(x):
    return x - 2**128 if x >= 2**127 else x


HARD = [0.0, 1.0, -1.0, 0.5, -0.5, 2.0**-64, -2.0**-64, 2.0**-70, -2.0**-70, 1e-18, -1e-18, -5e-17, -2.0**-54, -2.0**-53,
        1.0 - 2.0**-53, -(1.0 - 2.0**-53), 3.0 - 2.0**-51, -3.0 + 2.0**-51, 2.0**40 + 0.25, -(2.0**40 + 0.25), 2.0**61, -2.0**61,
        0.1, -0.1, 123.456, -123.456]


def test_split_is_trunc_times_2_64(lib):
    rng = random.Random(1)
    vals = HARD + [rng.choice((-1, 1)) * rng.random() * 2.0 ** rng.randint(-80, 40) for _ in range(4000)]
    for v in vals:
        assert words(lib, v) == exact(v), v


def test_sum_is_order_independent_and_exact(lib):
    rng = random.Random(2)
    for trial in range(50):
        n = rng.randint(1, 300)
        v = [rng.choice((-1, 1)) * rng.random() * 2.0 ** rng.randint(-70, 4) for _ in range(n)]
        if trial % 5 == 0:
            v += [-1e-18, -2.0**-54, 1.0 - 2.0**-53]
        ref = sum(exact(x) for x in v) % 2**128
        outs = []
        for perm in range(3):
            w = list(v)
            rng.shuffle(w)
            arr = (C.c_double * len(w))(*w)
            lo, hi = C.c_ulonglong(), C.c_ulonglong()
            lib.sum(arr, len(w), C.byref(lo), C.byref(hi))
            outs.append((hi.value << 64) | lo.value)
        assert outs[0] == outs[1] == outs[2] == ref
        got = lib.value(ref & (2**64 - 1), ref >> 64)
        true = Fraction(signed(ref), 2**64)
        assert abs(Fraction(got) - true) <= abs(true) * Fraction(2, 2**53) + Fraction(1, 2**64)
        # the truncated sum is within n * 2^-64 of the exact sum of the doubles
        assert abs(true - sum(Fraction(x) for x in v)) <= Fraction(len(v), 2**64)


def test_value_of_small_negative_totals(lib):
    """A total just below zero keeps its magnitude (no absorption into -1 + 1)."""
    for v in (-1e-18, -3.5e-17, -2.0**-60, -1e-10):
        x = exact(v)
        got = lib.value(x & (2**64 - 1), x >> 64)
        assert got != 0.0 and abs(got - v) <= 2.0**-64 + abs(v) * 2.0**-52, (v, got)
    assert math.copysign(1.0, lib.value(0, 0)) == 1.0

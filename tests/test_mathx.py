"""CPU: the backend's FP64 friction pow and boundary sine (ptha_amd/csrc/swh_mathx.h),
built for the host from the same header the kernels include.

Pins:
* correct rounding: on seeded arguments in the ranges the configurations reach, every
  result equals the exact value rounded to the nearest double, with the exact value
  from Python's decimal arithmetic at 50 digits (pow by decimal's power, sin by its
  Taylor series after reduction by a 50-digit pi);
* glibc: the results differ from the reference's libm only where glibc itself is not
  correctly rounded (rare), which bounds the ulp seeds of FP64 divergence;
* exact cases: powers of two, sin(0).
The sweep over 10^8 arguments against quad precision is tests/mathx_sweep.cpp.
"""
import ctypes
import math
import os
import shutil
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mathx_sweep.cpp")


@pytest.fixture(scope="module")
def mx(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("mathx")
    so = str(d / "libmathx.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", SRC, "-o", so], check=True)
    lib = ctypes.CDLL(so)
    P = ctypes.POINTER(ctypes.c_double)
    lib.mx_pow.argtypes = [ctypes.c_int, P, P, P]
    lib.mx_sin.argtypes = [ctypes.c_int, P, P]

    def p(a):
        return a.ctypes.data_as(P)

    def pw(x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
        out = np.empty_like(x)
        lib.mx_pow(len(x), p(x), p(y), p(out))
        return out

    def sn(t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        out = np.empty_like(t)
        lib.mx_sin(len(t), p(t), p(out))
        return out

    return pw, sn


def _pi50():
    getcontext().prec = 60

    def atan_inv(n):
        x = Decimal(1) / n
        s, t, k, sign = Decimal(0), x, 1, 1
        while t > Decimal(10) ** -65:
            s += sign * t / k
            t *= x * x
            k += 2
            sign = -sign
        return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)


def _sin_exact(t, pi):
    # reduce by 2 pi exactly in decimal, then Taylor series
    getcontext().prec = 60
    x = Decimal(t)
    two_pi = 2 * pi
    x = x - two_pi * (x / two_pi).to_integral_value()
    s, term, k = Decimal(0), x, 1
    while abs(term) > Decimal(10) ** -62:
        s += term
        term = -term * x * x / ((k + 1) * (k + 2))
        k += 2
    return float(s)


def test_pow_correctly_rounded_and_glibc(mx):
    pw, _ = mx
    rng = np.random.default_rng(20261203)
    x = np.exp(rng.uniform(math.log(1e-5), math.log(1e4), 20000))
    for y in (-7.0 / 3.0, -2.0):
        got = pw(x, y)
        getcontext().prec = 50
        exact = np.array([float(Decimal(v) ** Decimal(y)) for v in x])
        assert np.array_equal(got, exact), y
        glibc = np.array([math.pow(v, y) for v in x])
        assert (got != glibc).mean() < 5e-3, y


def test_sin_correctly_rounded_and_glibc(mx):
    _, sn = mx
    rng = np.random.default_rng(20261204)
    t = rng.uniform(-100.0, 100.0, 20000)
    pi = _pi50()
    got = sn(t)
    exact = np.array([_sin_exact(v, pi) for v in t])
    assert np.array_equal(got, exact)
    glibc = np.sin(t)
    assert (got != glibc).mean() < 5e-3


def test_exact_cases(mx):
    pw, sn = mx
    k = np.arange(-40, 41, dtype=np.float64)
    assert np.array_equal(pw(2.0 ** k, -2.0), 2.0 ** (-2 * k))
    assert np.array_equal(pw(np.array([512.0, 0.25, 1024.0]), -1.0), np.array([2.0 ** -9, 4.0, 2.0 ** -10]))
    assert sn(np.array([0.0]))[0] == 0.0
    # outside the handled range the library function is used
    assert math.isnan(pw(np.array([-1.0]), -7.0 / 3.0)[0])


def test_sweep_against_quad(tmp_path):
    """The sweep program on 10^6 arguments per function: no result differs from the
    correctly rounded one."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "sweep")
    r = subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fopenmp", "-DSWEEP_MAIN", SRC, "-lquadmath", "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("libquadmath not available: " + r.stderr[-200:])
    import json
    out = json.loads(subprocess.run([exe, "1000000"], check=True, capture_output=True, text=True).stdout)
    assert out["pow_m7o3"]["differs_from_correct_rounding"] == 0
    assert out["pow_m2"]["differs_from_correct_rounding"] == 0
    assert out["sin"]["differs_from_correct_rounding"] == 0
    # every difference from glibc is a glibc result that is not correctly rounded
    assert out["pow_m7o3"]["differs_from_glibc"] == out["pow_m7o3"]["glibc_differs_from_correct_rounding"]
    assert out["sin"]["differs_from_glibc"] == out["sin"]["glibc_differs_from_correct_rounding"]

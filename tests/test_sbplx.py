"""The library's resumable SBPLX (pgure-svt_amd/csrc/sbplx.hpp, replayed one
evaluation at a time by the lockstep search) against the oracle's
straight-through restatement (oracle/pgsvt_oracle.cpp, namespace sbplx):
the same sequence of evaluated lambdas, return code and last lambda on
objectives of the shapes the PGURE search meets (interior minimum, minimum at
or beyond a bound, a kink)."""

import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


@pytest.fixture(scope="module")
def product(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    so = str(tmp_path_factory.mktemp("sbplx") / "libsbplx_driver.so")
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "sbplx_driver.cpp")],
                   check=True)
    L = C.CDLL(so)
    L.product_sbplx.argtypes = [FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.product_sbplx.restype = C.c_int
    return L


def run_product(L, f, start, lb, ub, step, ftol=1e-7, xtol=1e-12, maxeval=500):
    xs = []

    def g(x, _ctx):
        xs.append(x)
        return float(f(x))

    cb = FN(g)
    xl, ne = C.c_double(), C.c_int()
    rc = L.product_sbplx(cb, None, start, lb, ub, step, ftol, xtol, maxeval, C.byref(xl), C.byref(ne))
    return rc, xs, xl.value, ne.value


def run_oracle(f, start, lb, ub, step, ftol=1e-7, xtol=1e-12, maxeval=500):
    xs = []

    def g(x):
        xs.append(x)
        return float(f(x))

    rc, xb, fb, xl, ne = O.sbplx(g, start, lb, ub, step, ftol, xtol, maxeval)
    return rc, xs, xl, ne


CASES = [
    ("interior", lambda x: (x - 0.25) ** 2 * (1 + 0.3 * x) + 1.0, 0.12),
    ("at_lower_bound", lambda x: (x + 0.5) ** 2, 0.3),
    ("beyond_upper_bound", lambda x: (x - 500.0) ** 2, 2.0),
    ("kink", lambda x: abs(x - 0.37) + 2.0, 0.11),
    ("flat_then_rising", lambda x: max(x - 0.8, 0.0) ** 2 + 1.0, 0.2),
]


@pytest.mark.parametrize("name,f,start", CASES, ids=[c[0] for c in CASES])
def test_product_search_matches_oracle(product, name, f, start):
    ub = max(100.0, start)
    rp, xp, lp, nep = run_product(product, f, start, 0.0, ub, np.sqrt(start))
    ro, xo, lo, neo = run_oracle(f, start, 0.0, ub, np.sqrt(start))
    assert xp == xo, name
    assert rp == ro and lp == lo and nep == len(xp) == len(xo)


def test_evaluation_cap(product):
    f = CASES[0][1]
    rp, xp, lp, nep = run_product(product, f, 0.12, 0.0, 100.0, np.sqrt(0.12), ftol=0.0, xtol=0.0, maxeval=7)
    ro, xo, lo, neo = run_oracle(f, 0.12, 0.0, 100.0, np.sqrt(0.12), ftol=0.0, xtol=0.0, maxeval=7)
    assert xp == xo and len(xp) == 7 and rp == ro == 5 and lp == lo

"""Accuracy of the galactic-binary generator's sine/cosine (csrc/flr_sincos.hpp:phase_sincos),
compiled for the host with g++ from the same header the device code includes (no GPU needed).

Pinned against the C library's sin/cos (Python's math module, within an ulp of the exact
values) on phases of the sizes the generator meets (up to a few 1e7 rad for year-long
sources), on arguments next to multiples of pi/2 (where the reduction cancels), on every
quadrant of both signs, and on the library fallback above 2^28. A dropped reduction term
(pi/2 - P1 ~ 6e-17 times n ~ 1e7) or a wrong quadrant rotation fails these bounds.
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "lisa-on-gpu_amd", "csrc")

SHIM = r"""
#include "flr_sincos.hpp"
extern "C" void phase_sincos_many(const double *x, long n, double *s, double *c)
{
    for (long i = 0; i < n; ++i) flr::phase_sincos(x[i], s + i, c + i);
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("sincos")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "libshim.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-DFLR_HOST_DEVICE=",
                    "-I", CSRC, str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.phase_sincos_many.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                                    ctypes.c_void_p]
    return L


def _run(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    lib.phase_sincos_many(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
    return s, c


def _libm(x):
    return (np.array([math.sin(v) for v in x]), np.array([math.cos(v) for v in x]))


def _check(lib, x, tol=3e-16):
    s, c = _run(lib, x)
    es, ec = _libm(x)
    assert np.abs(s - es).max() <= tol, float(np.abs(s - es).max())
    assert np.abs(c - ec).max() <= tol, float(np.abs(c - ec).max())


def test_generator_phases(lib):
    rng = np.random.default_rng(7)
    # 2 pi (f t + fdot t^2 / 2) for f in [1e-4, 1e-2] Hz, t up to 4.2e7 s, either sign
    f = 10 ** rng.uniform(-4, -2, 40000)
    t = rng.uniform(0, 4.2e7, 40000)
    x = 2 * np.pi * f * t * rng.choice([-1.0, 1.0], 40000)
    _check(lib, x)
    _check(lib, rng.uniform(-4, 4, 20000))
    _check(lib, 10 ** rng.uniform(-300, 8.4, 20000) * rng.choice([-1.0, 1.0], 20000))


def test_next_to_multiples_of_half_pi(lib):
    rng = np.random.default_rng(11)
    n = rng.integers(-(1 << 27), 1 << 27, 20000).astype(np.float64)
    base = n * (math.pi / 2)
    x = base + rng.uniform(-1e-6, 1e-6, base.size)
    _check(lib, x)
    # small results stay small: absolute error, not relative, is what h needs
    s, c = _run(lib, np.array([math.pi, 2 * math.pi, 1e7 * math.pi]))
    assert abs(s[0] - math.sin(math.pi)) <= 1e-18 and abs(s[1] - math.sin(2 * math.pi)) <= 1e-18


def test_each_quadrant_and_special_values(lib):
    q = np.arange(-8, 9)
    x = q * (math.pi / 2) + 0.3
    _check(lib, x)
    s, c = _run(lib, np.array([0.0, -0.0]))
    assert s[0] == 0.0 and c[0] == 1.0 and c[1] == 1.0
    # above 2^28 the library's own reduction is used
    big = np.array([2.0 ** 28, -3.3e9, 1.0e15])
    _check(lib, big, tol=2.3e-16)
    s, c = _run(lib, np.array([np.nan, np.inf]))
    assert np.isnan(s).all() and np.isnan(c).all()

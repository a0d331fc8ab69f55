"""The closed form for n repeated floating-point additions (sfp_repeat_add,
used by the step to push a photon several steps at once) against the plain
loop of n additions, bit for bit, on the host build of the same header.

Cases: magnitudes across many binades, both signs, binade and zero
crossings, increments that are exact ties at the current ulp (both parities),
increments too small to move x, zeros of both signs, subnormals and
non-finite values.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <stdint.h>
#include "sfp_math.h"
void closed(const double* x, const double* h, const uint32_t* n, double* out, int64_t cnt) {
  for (int64_t i = 0; i < cnt; ++i) out[i] = sfp_repeat_add(x[i], h[i], n[i]);
}
void naive(const double* x, const double* h, const uint32_t* n, double* out, int64_t cnt) {
  for (int64_t i = 0; i < cnt; ++i) {
    volatile double v = x[i];
    for (uint32_t k = 0; k < n[i]; ++k) v = v + h[i];
    out[i] = v;
  }
}
"""


@pytest.fixture(scope="module")
def lib():
    d = tempfile.mkdtemp(prefix="sfp_repeat_add_")
    c = os.path.join(d, "t.c")
    so = os.path.join(d, "t.so")
    with open(c, "w") as f:
        f.write(SRC)
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                           "-I", os.path.join(ROOT, "sfparticles_amd", "csrc"), "-o", so, c])
    L = C.CDLL(so)
    for fn in (L.closed, L.naive):
        fn.argtypes = [C.c_void_p] * 4 + [C.c_int64]
    return L


def _run(L, x, h, n):
    x = np.ascontiguousarray(x, np.float64)
    h = np.ascontiguousarray(h, np.float64)
    n = np.ascontiguousarray(n, np.uint32)
    a = np.empty_like(x)
    b = np.empty_like(x)
    L.closed(x.ctypes.data, h.ctypes.data, n.ctypes.data, a.ctypes.data, x.size)
    L.naive(x.ctypes.data, h.ctypes.data, n.ctypes.data, b.ctypes.data, x.size)
    return a, b


def _check(L, x, h, n):
    a, b = _run(L, x, h, n)
    same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, [(x[i], h[i], int(n[i]), a[i], b[i]) for i in bad[:5]]


def test_random_magnitudes(lib):
    rng = np.random.Generator(np.random.PCG64(1))
    m = 20000
    x = rng.choice([-1.0, 1.0], m) * 10.0 ** rng.uniform(-12, -2, m)
    h = rng.choice([-1.0, 1.0], m) * np.abs(x) * 10.0 ** rng.uniform(-17, -1, m)
    n = rng.integers(0, 3000, m)
    _check(lib, x, h, n)


def test_photon_like_positions(lib):
    """The step's case: positions of micrometres, increments c dt/2 * u/|u|."""
    rng = np.random.Generator(np.random.PCG64(2))
    m = 20000
    x = rng.normal(0, 10e-6, m)
    cdt_half = 299792458.0 * 2.6685127615e-17 / 2
    ux = rng.normal(0, 1, m)
    ig = 1 / np.sqrt(ux ** 2 + rng.normal(0, 1, m) ** 2 + 1)
    h = cdt_half * ig * ux
    n = rng.integers(1, 6000, m)
    _check(lib, x, h, n)


def test_zero_crossings_and_tiny(lib):
    rng = np.random.Generator(np.random.PCG64(3))
    m = 5000
    x = rng.choice([-1.0, 1.0], m) * 10.0 ** rng.uniform(-300, -8, m)
    h = -np.sign(x) * 10.0 ** rng.uniform(-12, -9, m)
    n = rng.integers(1, 2000, m)
    _check(lib, x, h, n)
    # subnormal start and increment
    x2 = np.array([5e-324, -5e-324, 1e-310, -1e-310, 2.2250738585072014e-308] * 20)
    h2 = np.array([5e-324, 1e-320, -3e-321, 1e-309, -5e-324] * 20)
    n2 = np.arange(x2.size) * 37 % 500
    _check(lib, x2, h2, n2)


def test_binade_edges(lib):
    rng = np.random.Generator(np.random.PCG64(4))
    m = 8000
    e = rng.integers(-40, 10, m).astype(np.float64)
    ulps = rng.integers(-20, 20, m)
    x = np.ldexp(1.0, e.astype(int))
    x = np.nextafter(x, np.where(ulps > 0, np.inf, -np.inf))
    for _ in range(3):
        x = np.where(np.abs(ulps) > 1, np.nextafter(x, np.where(ulps > 0, np.inf, -np.inf)), x)
    x *= rng.choice([-1.0, 1.0], m)
    u = np.spacing(np.abs(x))
    h = rng.choice([-1.0, 1.0], m) * u * rng.uniform(0.01, 7.0, m)
    n = rng.integers(1, 400, m)
    _check(lib, x, h, n)


def test_ties(lib):
    """h an exact half-integer number of ulps of x: the rounding depends on
    the parity of x's last bit."""
    rng = np.random.Generator(np.random.PCG64(5))
    m = 8000
    x = rng.choice([-1.0, 1.0], m) * 10.0 ** rng.uniform(-9, -3, m)
    # odd and even last bits
    x = np.where(rng.random(m) < 0.5, np.nextafter(x, np.inf), x)
    u = np.spacing(np.abs(x))
    j = rng.integers(-6, 6, m)
    h = (j + 0.5) * u
    n = rng.integers(1, 3000, m)
    _check(lib, x, h, n)
    # ties at half an ulp (q = 0 or +-1 by parity) and at binade bottoms
    x2 = np.ldexp(1.0, rng.integers(-30, 0, 2000))
    x2 = np.where(rng.random(2000) < 0.5, np.nextafter(x2, np.inf), x2)
    u2 = np.spacing(x2)
    h2 = rng.choice([-0.5, 0.5, -1.5, 1.5, -0.25, 0.25], 2000) * u2
    _check(lib, x2, h2, rng.integers(1, 500, 2000))


def test_special_values(lib):
    x = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1.0, 1e300, -1e300,
                  1.7976931348623157e308, 3.0])
    h = np.array([0.0, 0.0, -0.0, -0.0, 0.0, -0.0, 1.0, np.inf, 1.0, np.inf, 1e300, -1e299,
                  1e292, np.nan])
    for n in (0, 1, 2, 7, 1000):
        _check(lib, x, h, np.full(x.size, n))

"""The closed form for n repeated floating-point additions
(sfp_repeat_add_closed, used by the step, through sfp_repeat_add, to push a
photon several steps at once) against the plain loop of n additions, bit for
bit, on the host build of the same header.

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
  for (int64_t i = 0; i < cnt; ++i) out[i] = sfp_repeat_add_closed(x[i], h[i], n[i]);
}
void public_(const double* x, const double* h, const uint32_t* n, double* out, int64_t cnt) {
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
    for fn in (L.closed, L.naive, L.public_):
        fn.argtypes = [C.c_void_p] * 4 + [C.c_int64]
    return L


def _run(L, x, h, n):
    x = np.ascontiguousarray(x, np.float64)
    h = np.ascontiguousarray(h, np.float64)
    n = np.ascontiguousarray(n, np.uint32)
    a = np.empty_like(x)
    b = np.empty_like(x)
    c = np.empty_like(x)
    L.closed(x.ctypes.data, h.ctypes.data, n.ctypes.data, a.ctypes.data, x.size)
    L.naive(x.ctypes.data, h.ctypes.data, n.ctypes.data, b.ctypes.data, x.size)
    L.public_(x.ctypes.data, h.ctypes.data, n.ctypes.data, c.ctypes.data, x.size)
    return a, b, c


def _check(L, x, h, n):
    """The closed form and the public entry (one by one up to 64 additions,
    the closed form beyond) against the plain loop."""
    a, b, c = _run(L, x, h, n)
    for r in (a, c):
        same = (r.view(np.uint64) == b.view(np.uint64)) | (np.isnan(r) & np.isnan(b))
        bad = np.nonzero(~same)[0]
        assert bad.size == 0, [(x[i], h[i], int(n[i]), r[i], b[i]) for i in bad[:5]]


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


def test_steps_that_just_fit_in_the_binade(lib):
    """Integer increments of q ulps for which n q equals the room left in the
    binade, exceeds it by one, or falls one short: the test that all n steps
    stay inside the binade, with q below and above 2^32 ulps (n q = room + 1
    lands exactly on the next binade's end, which is still exact)."""
    rng = np.random.Generator(np.random.PCG64(6))
    xs, hs, ns = [], [], []
    for _ in range(3000):
        q = int(rng.choice([1, 3, 1000, 2 ** 20 + 7, 2 ** 33 + 5, 2 ** 40 + 3]))
        n = int(rng.integers(1, 3000))
        if n * q >= 2 ** 52 - 2:
            n = (2 ** 52 - 2) // q
        slack = int(rng.integers(-3, 2))
        e = int(rng.integers(-30, 0))
        if rng.random() < 0.5:  # upwards: room = 2^53 - 1 - k
            k = 2 ** 53 - 1 - n * q - slack
            sign = 1.0
        else:  # downwards: room = k - 2^52 - 1
            k = 2 ** 52 + 1 + n * q + slack
            sign = -1.0
        k = min(max(k, 2 ** 52), 2 ** 53 - 1)
        xs.append(np.ldexp(float(k), e - 52))
        hs.append(sign * np.ldexp(float(q), e - 52))
        ns.append(n)
    x, h, n = np.array(xs), np.array(hs), np.array(ns)
    flip = rng.random(x.size) < 0.5  # negative x mirrors the direction
    _check(lib, np.where(flip, -x, x), np.where(flip, -h, h), n)


def test_binade_bottom_small_decrements(lib):
    """x exactly a power of two and h a decrement of less than half an ulp of
    x's binade: the first addition lands in the binade below, where the ulp
    is half as large, unless |h| is at most a quarter of x's ulp."""
    rng = np.random.Generator(np.random.PCG64(7))
    m = 4000
    x = rng.choice([-1.0, 1.0], m) * np.ldexp(1.0, rng.integers(-40, 10, m))
    u = np.spacing(np.abs(x))
    frac = rng.choice([0.1, 0.25, 0.3, 0.45, 0.49, 0.5, 0.7], m)
    h = -np.sign(x) * frac * u
    _check(lib, x, h, rng.integers(1, 300, m))

"""The device's correctly rounded acos and sincos (gpluto_dev_amd/csrc/crmath.hpp),
built for the host, against the host libm that the reference's CPU build calls
(Src/RMHD/rmhd.hpp:441-444) and against exact values from mpmath.

The RMHD quartic's resolvent cubic takes theta = acos(arg)/3 with arg in
[-1, 1], then cos(theta) and sin(theta).  glibc 2.35 returns the correctly
rounded double for all but about 7e-4 (acos) and 3e-3 (sin or cos of one
argument) of the arguments; the device library is faithful only.  The correctly rounded
routines therefore give the device glibc's bits except where glibc itself is
off by one rounding, and every such disagreement inspected here is glibc's
rounding, not ours (mpmath decides)."""
import ctypes
import math
import os
from fractions import Fraction

import mpmath
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhost_crmath.so")
THREADS = min(8, os.cpu_count() or 1)

ACOS_UNIFORM, ACOS_NEAR_ONE, SINCOS_UNIFORM, SINCOS_CUBIC, ACOS_PERTURBED_SEED = range(5)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "host")], check=True)
    L = ctypes.CDLL(LIB)
    L.crm_acos.restype = ctypes.c_double
    L.crm_acos.argtypes = [ctypes.c_double]
    L.crm_sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double)]
    L.crm_compare.restype = ctypes.c_longlong
    L.crm_compare.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_ulonglong, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                              ctypes.c_int]
    return L


def _compare(L, mode, n, lo=0.0, hi=0.0, seed=1, nfirst=16):
    first = (ctypes.c_double * nfirst)()
    c = L.crm_compare(mode, n, seed, lo, hi, THREADS, first, nfirst)
    return c, [first[i] for i in range(min(c, nfirst))]


def _nearest(v) -> float:
    sign, man, exp, _ = mpmath.mpf(v)._mpf_
    f = Fraction(int(man)) * Fraction(2) ** int(exp)
    return float(-f if sign else f)


def _sincos(L, x):
    s, c = ctypes.c_double(), ctypes.c_double()
    L.crm_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def test_special_values(lib):
    for x in (1.0, -1.0, 0.0, -0.0, 0.5, -0.5, 5e-324, 1 - 2 ** -53, -1 + 2 ** -53):
        assert lib.crm_acos(x) == math.acos(x), x
    assert math.isnan(lib.crm_acos(float("nan")))
    assert _sincos(lib, 0.0) == (0.0, 1.0)
    s, c = _sincos(lib, float("nan"))
    assert math.isnan(s) and math.isnan(c)


def test_correctly_rounded_against_mpmath(lib):
    """20000 arguments each: acos on [-1, 1] and near +-1, sin/cos on the
    cubic's [0, pi/3] and the acos Newton step's [0, 2.19]: every result is the
    double nearest the exact value."""
    mpmath.mp.dps = 50
    g = __import__("numpy").random.default_rng(7)
    xs = list(g.uniform(-1, 1, 10000)) + [(1 - 2.0 ** -e) * s for e, s in
                                          zip(g.integers(1, 53, 10000), g.choice([-1, 1], 10000))]
    for x in xs:
        assert lib.crm_acos(float(x)) == _nearest(mpmath.acos(mpmath.mpf(float(x)))), x
    for x in list(g.uniform(0, math.pi / 3, 10000)) + list(g.uniform(0, 2.19, 10000)):
        s, c = _sincos(lib, float(x))
        assert s == _nearest(mpmath.sin(mpmath.mpf(float(x)))), x
        assert c == _nearest(mpmath.cos(mpmath.mpf(float(x)))), x


def _glibc_is_the_one_off(fn, x, ours, theirs):
    exact = fn(mpmath.mpf(x))
    assert ours == _nearest(exact), (x, ours)
    assert abs(mpmath.mpf(theirs) - exact) > abs(mpmath.mpf(ours) - exact), (x, theirs)


# (mode, arguments, bound on the mismatch fraction) -- the counts measured for
# seed 1 are recorded with the design notes
@pytest.mark.parametrize("mode,n,lo,hi,bound", [
    (ACOS_UNIFORM, 10 ** 8, -1.0, 1.0, 1e-3),
    (ACOS_NEAR_ONE, 2 * 10 ** 7, 0.0, 0.0, 1e-3),
    (SINCOS_UNIFORM, 10 ** 8, 0.0, math.pi / 3, 4e-3),
    (SINCOS_CUBIC, 10 ** 8, 0.0, 0.0, 4e-3),
])
def test_against_host_libm(lib, mode, n, lo, hi, bound):
    mpmath.mp.dps = 50
    count, first = _compare(lib, mode, n, lo, hi)
    print(f"mode {mode}: {count} of {n} arguments differ from glibc ({count / n:.2e})")
    assert count <= bound * n
    for x in first:
        if mode in (ACOS_UNIFORM, ACOS_NEAR_ONE):
            _glibc_is_the_one_off(mpmath.acos, x, lib.crm_acos(x), math.acos(x))
        else:
            s, c = _sincos(lib, x)
            if s != math.sin(x):
                _glibc_is_the_one_off(mpmath.sin, x, s, math.sin(x))
            if c != math.cos(x):
                _glibc_is_the_one_off(mpmath.cos, x, c, math.cos(x))


def test_newton_step_corrects_a_seed_a_few_ulp_off(lib):
    """The device's acos and asin seeds are a few ulp from the exact value:
    seeds moved 1 to 4 ulp off give exactly the results of the exact seeds."""
    count, first = _compare(lib, ACOS_PERTURBED_SEED, 2 * 10 ** 7, -1.0, 1.0)
    assert count == 0, first

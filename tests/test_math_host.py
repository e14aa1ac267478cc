"""The push's hypot (trench01_amd/csrc/th_math.h) compiled for the host and
checked against glibc's hypot, the function NumPy's np.hypot calls (through
the oracle's binding), on random operands over the whole exponent range and on
the special cases (zeros, subnormals, infinities, NaN, widely separated
operands, the rescaling thresholds).  The device self-test repeats the random
part on the GPU."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import trace_ref

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def host_hypot(tmp_path_factory):
    d = tmp_path_factory.mktemp("thmath")
    src = d / "h.c"
    src.write_text('#include "%s"\nvoid th_hypot_arr(const double *a, const double *b, double *o, long n) '
                   '{ for (long k = 0; k < n; ++k) o[k] = th_hypot(a[k], b[k]); }\n'
                   'void th_hypot2_arr(const double *a, const double *b, const double *c, const double *d, '
                   'double *o0, double *o1, long n) '
                   '{ for (long k = 0; k < n; ++k) th_hypot2(a[k], b[k], c[k], d[k], &o0[k], &o1[k]); }\n'
                   % (ROOT / "trench01_amd" / "csrc" / "th_math.h"))
    lib = d / "libh.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", str(src), "-o", str(lib),
                    "-lm"], check=True)
    L = ctypes.CDLL(str(lib))
    dp = ctypes.POINTER(ctypes.c_double)
    L.th_hypot_arr.argtypes = [dp, dp, dp, ctypes.c_long]

    def f(a, b):
        a, b = np.ascontiguousarray(a, float), np.ascontiguousarray(b, float)
        o = np.empty_like(a)
        L.th_hypot_arr(a.ctypes.data_as(dp), b.ctypes.data_as(dp), o.ctypes.data_as(dp), a.size)
        return o
    f.lib = L
    return f


def same(x, y):
    return np.array_equal(x, y, equal_nan=True) and np.array_equal(np.signbit(x), np.signbit(y))


def test_random_operands_over_the_exponent_range(host_hypot):
    rng = np.random.default_rng(0)
    n = 1 << 20
    ea = rng.integers(-1074, 1023, n)
    eb = ea + rng.integers(-60, 61, n)
    a = np.ldexp(rng.random(n) + 0.5, ea) * rng.choice([-1, 1], n)
    b = np.ldexp(rng.random(n) + 0.5, np.clip(eb, -1074, 1023)) * rng.choice([-1, 1], n)
    assert same(host_hypot(a, b), trace_ref.glibc_hypot(a, b))


def test_push_range_operands(host_hypot):
    """Velocities and accelerations of the push: 1e2..1e8 m/s, 1e10..1e22 m/s^2."""
    rng = np.random.default_rng(1)
    n = 1 << 20
    for lo, hi in ((2, 8), (10, 22)):
        a = 10.0 ** rng.uniform(lo, hi, n) * rng.choice([-1, 1], n)
        b = 10.0 ** rng.uniform(lo, hi, n) * rng.choice([-1, 1], n)
        assert same(host_hypot(a, b), trace_ref.glibc_hypot(a, b))


def test_widely_separated_operands_in_the_core_range(host_hypot):
    """ay <= ax 2^-54 inside the core's exponent range goes through the core (no ratio test): it must return
    ax, as glibc's widely-separated branch does."""
    rng = np.random.default_rng(2)
    n = 1 << 20
    ax = np.ldexp(rng.random(n) + 0.5, rng.integers(-330, 511, n))
    ay = np.ldexp(ax * (rng.random(n) + 0.5), -rng.integers(54, 60, n))
    ay = np.where(ay >= 2.0 ** -383, ay, 2.0 ** -383)
    for a, b in ((ax, ay), (ay, -ax), (-ax, ay)):
        assert same(host_hypot(a, b), trace_ref.glibc_hypot(a, b))


def test_special_values(host_hypot):
    vals = np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 2.0 ** -459, 2.0 ** -460, 2.0 ** 511,
                     2.0 ** 512, 1.7976931348623157e308, np.inf, -np.inf, np.nan, 1.0, 3.0, 4.0,
                     1e-300, 1e300, 2.0 ** -54, 2.0 ** 54])
    a, b = np.meshgrid(vals, vals)
    a, b = a.ravel(), b.ravel()
    assert same(host_hypot(a, b), trace_ref.glibc_hypot(a, b))
    # widely separated pairs around the 2^-54 ratio
    x = np.ldexp(1.0, np.arange(-40, 40))
    for r in (2.0 ** -53, 2.0 ** -54, 2.0 ** -55):
        y = x * r * (1 + np.linspace(-1e-3, 1e-3, x.size))
        assert same(host_hypot(x, y), trace_ref.glibc_hypot(x, y))


def test_paired_hypot_is_two_hypots(host_hypot):
    """th_hypot2 (both cores evaluated, out-of-range pairs recomputed) gives
    the doubles of two th_hypot calls, glibc's, for every combination of a
    pair inside and outside the core's range (special values included)."""
    L = host_hypot.lib
    dp = ctypes.POINTER(ctypes.c_double)
    L.th_hypot2_arr.argtypes = [dp, dp, dp, dp, dp, dp, ctypes.c_long]
    rng = np.random.default_rng(5)
    n = 1 << 18
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 2.0 ** -383, 2.0 ** -384, 2.0 ** 511,
                        2.0 ** 512, 1.0, -3.5e7])
    def draw():
        v = np.ldexp(rng.random(n) + 0.5, rng.integers(-1074, 1023, n)) * rng.choice([-1, 1], n)
        k = rng.random(n) < 0.1
        v[k] = rng.choice(special, k.sum())
        return v
    a, b, c, d = (np.ascontiguousarray(draw()) for _ in range(4))
    o0, o1 = np.empty(n), np.empty(n)
    L.th_hypot2_arr(*(x.ctypes.data_as(dp) for x in (a, b, c, d, o0, o1)), n)
    assert same(o0, trace_ref.glibc_hypot(a, b))
    assert same(o1, trace_ref.glibc_hypot(c, d))

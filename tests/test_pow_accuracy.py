"""Accuracy of the chain's own log / exp / pow (mhm_amd/csrc/chain.h).

The cell kernels evaluate the reference's x**y (mo_canopy_interc.f90:121,
mo_soil_moisture.f90:219, mo_runoff.f90:134) as exp(y log x) with a log and
an exp written from their series (no library calls), so the same source
compiles for the host and for gfx950.  Pinned here on the host build:
  * against mpmath at 200 bits: log within 1.25 ulp (the table method's
    near-1 cancellation reaches 1.0002), exp within 1 ulp, pow within
    (2 + 3 |y log x|) ulp (log's 1 ulp and the rounding of y log x, amplified
    by exp), over the ranges the chain feeds them and beyond;
  * closed forms: log(2^k) = k ln 2, log(1) = 0, exp(0) = 1, exp(1) = e,
    exp(z) exp(-z) = 1, x**1 = x, 4**0.5 = 2, 2**10 = 1024, overflow and
    underflow, subnormal arguments;
  * planted mistakes (a dropped series term in exact-rounding Python copies
    of log_pos and exp_fin) fail the 1-ulp checks.
The GPU test checks that the device build gives the host's bits.  The
canopy's x**(2/3) is pow23 (a cube root, no log and exp): within 3 ulp on
the host and on the device (whose seed comes from hardware single-precision
log2 / exp2, so the two may differ in the last bit), exact on exact cubes.
"""
import ctypes as C
import math
import os
from fractions import Fraction

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "native", "libmhmcheck.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from mhm_amd import build
        build.build_check()
    h = C.CDLL(LIB)
    for f in (h.mhmhip_check_eval_host, h.mhmhip_check_eval_device):
        f.restype = C.c_int
        f.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 3 + [C.c_int64]
    return h


def _eval(lib, which, x, y=None, device=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    f = lib.mhmhip_check_eval_device if device else lib.mhmhip_check_eval_host
    assert f(which, p(x), p(y), p(out), x.shape[0]) == 0
    return out


def _ulp_err(got, exact):
    """|got - exact| in units of the last place of the double nearest exact."""
    e = float(exact)
    return float(abs(mp.mpf(got) - exact) / mp.mpf(math.ulp(e)))


mp.mp.prec = 200


def test_log_within_a_quarter_more_than_one_ulp(lib):
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 3000)),
                        1.0 + rng.uniform(-0.3, 0.45, 3000),          # the reduction's edges
                        rng.uniform(1e-12, 2.0, 2000),
                        np.array([0.7071067811865475, 0.7071067811865476, 1.4142135623730951,
                                  1.0 - 2**-53, 1.0 + 2**-52, 5e-324, 2.2250738585072014e-308])])
    got = _eval(lib, 0, x)
    worst = max(_ulp_err(g, mp.log(mp.mpf(v))) for g, v in zip(got, x) if v != 1.0)
    assert worst <= 1.25, worst


def test_exp_within_one_ulp(lib):
    rng = np.random.default_rng(2)
    z = np.concatenate([rng.uniform(-700, 700, 3000), rng.uniform(-1, 1, 3000),
                        rng.uniform(-1e-8, 1e-8, 500), np.array([0.34657359, -0.34657359])])
    got = _eval(lib, 1, z)
    worst = max(_ulp_err(g, mp.exp(mp.mpf(v))) for g, v in zip(got, z))
    assert worst <= 1.0, worst


def test_pow_error_bound(lib):
    """x in the chain's ranges: ratios in (0, 1] (canopy, soil), storages up
    to 1e4 mm (unsaturated zone); exponents 2/3, 0.1..10."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(1e-9, 1.0, 4000), np.exp(rng.uniform(-20, 9.2, 4000))])
    y = np.concatenate([np.full(1000, 2.0 / 3.0), rng.uniform(0.1, 10.0, 7000)])
    got = _eval(lib, 2, x, y)
    for g, a, b in zip(got, x, y):
        exact = mp.power(mp.mpf(a), mp.mpf(b))
        bound = 2.0 + 3.0 * abs(b * math.log(a))
        assert _ulp_err(g, exact) <= bound, (a, b, g)


def test_batched_equals_scalar(lib):
    """exp_ylogx_v<4> (four powers in lockstep) gives each element the bits
    of the scalar evaluation."""
    rng = np.random.default_rng(6)
    x = np.exp(rng.uniform(-30, 9, 4001))
    y = rng.uniform(0.1, 10.0, 4001)
    assert np.array_equal(_eval(lib, 3, x, y), _eval(lib, 2, x, y))


def test_division_by_known_reciprocal_is_exact(lib):
    """div_rcp (x * RN(1/d) plus one fma correction) gives the bits of x / d
    for the divisors the kernels use it with (steps per day, seconds per
    step, monthly evaporation coefficients) and random ones, including
    quotients next to exact multiples."""
    rng = np.random.default_rng(7)
    d = np.concatenate([np.full(20000, 24.0), np.full(20000, 3600.0), np.full(20000, 86400.0),
                        rng.uniform(0.3, 2.0, 40000), np.exp(rng.uniform(-7, 7, 40000))])
    x = np.ldexp(rng.uniform(0.5, 1.0, d.shape[0]), rng.integers(-60, 60, d.shape[0]))
    x[:3000] = d[:3000] * rng.integers(1, 500, 3000)
    x[3000:6000] = np.nextafter(d[3000:6000] * rng.integers(1, 500, 3000), np.inf)
    assert np.array_equal(_eval(lib, 4, x, d), x / d)
    assert np.array_equal(_eval(lib, 5, x, d), x / d)


def test_closed_forms(lib):
    k = np.arange(-1074, 1024, dtype=np.float64)
    lg = _eval(lib, 0, np.ldexp(1.0, k.astype(int)))
    exact = [mp.mpf(int(v)) * mp.log(2) for v in k]
    assert max(_ulp_err(g, e) for g, e in zip(lg, exact) if e != 0) <= 0.5
    assert _eval(lib, 0, [1.0])[0] == 0.0
    assert _eval(lib, 1, [0.0])[0] == 1.0
    assert _ulp_err(_eval(lib, 1, [1.0])[0], mp.e) <= 0.5
    z = np.random.default_rng(5).uniform(-300, 300, 2000)
    prod = _eval(lib, 1, z) * _eval(lib, 1, -z)
    assert np.max(np.abs(prod - 1.0)) <= 4.5e-16
    x = np.array([3.7, 0.123, 1e-5, 42.0])
    assert np.all(np.abs(_eval(lib, 2, x, np.ones(4)) - x) <= (2 + 3 * np.abs(np.log(x))) * np.spacing(x))
    assert _eval(lib, 2, [4.0], [0.5])[0] == 2.0
    assert _eval(lib, 2, [2.0], [10.0])[0] == 1024.0
    assert _eval(lib, 1, [800.0])[0] == np.inf
    assert _eval(lib, 1, [-800.0])[0] == 0.0


LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _table(name):
    """The log rows (4 values each) or the exp rows (2 each) of mathtab.h's
    kMathTab."""
    import re
    path = os.path.join(ROOT, "mhm_amd", "csrc", "mathtab.h")
    text = open(path).read()
    n_log = int(re.search(r"kLogTabN = (\d+);", text).group(1))
    body = text[text.index("kMathTab["):]
    body = body[body.index("{") + 1:body.index("};")]
    body = re.sub(r"//[^\n]*", "", body)
    vals = [float.fromhex(v) if "x" in v else float(v)
            for v in re.findall(r"-?0x[0-9a-fp.+-]+|0\.0", body)]
    if name == "log":
        return [vals[4 * k:4 * k + 4] for k in range(n_log)]
    e = vals[4 * n_log:]
    return [e[2 * k:2 * k + 2] for k in range(len(e) // 2)]


def _log_copy(x, degree=7):
    """log_pos restated in Python (every operation correctly rounded)."""
    tab = _table("log")
    m, e = math.frexp(x)
    jd = float(round(m * 256.0))
    f = _fma(jd, -0.00390625, m)
    inv, hi, lo, _ = tab[int(jd) - 128]
    u = f * inv
    p = 1.0 / degree * (1 if degree % 2 else -1)
    for k in range(degree - 1, 1, -1):
        p = _fma(p, u, (1.0 if k % 2 else -1.0) / k)
    poly = _fma(u * u, p, u)
    return _fma(float(e), LN2_HI, hi) + (poly + _fma(float(e), LN2_LO, lo))


def _exp_copy(z, top=4):
    """exp_fin restated in Python; top: the last Taylor term of exp(r) - 1."""
    tab = _table("exp")
    n = float(round(z * float.fromhex("0x1.71547652b82fep+8")))
    r = _fma(-n, LN2_HI / 256.0, z)
    r = _fma(-n, LN2_LO / 256.0, r)
    q = 1.0 / math.factorial(top)
    for k in range(top - 1, 1, -1):
        q = _fma(q, r, 1.0 / math.factorial(k))
    pm = _fma(r * r, q, r)
    hi, lo = tab[int(n) & 255]
    return math.ldexp(hi + _fma(hi, pm, lo), int(n) >> 8)


def test_copies_match_the_library_and_planted_mistakes_fail(lib):
    xs = [0.7071067811865476, 1.4142135623730949, 1.41, 0.7072, 3.3e-7, 812.5, 0.70718, 1.0 + 1.0 / 256]
    zs = [0.3465, -0.3465, 5.0 + 0.3465, -7.0 - 0.3465]
    assert [_log_copy(x) for x in xs] == list(_eval(lib, 0, xs))
    assert [_exp_copy(z) for z in zs] == list(_eval(lib, 1, zs))
    # a dropped term: log1p through u^6, exp(r) - 1 through r^3
    assert max(_ulp_err(_log_copy(x, degree=6), mp.log(mp.mpf(x))) for x in xs) > 1.25
    assert max(_ulp_err(_exp_copy(z, top=3), mp.exp(mp.mpf(z))) for z in zs) > 1.0


def _pow23_sample(rng, n):
    return np.concatenate([np.exp(rng.uniform(-700, 0, n)), rng.uniform(0.0, 1.0, n),
                           np.exp(rng.uniform(0, 700, n // 4)),
                           np.array([5e-324, 2.2250738585072014e-308, 2**-1074 * 3, 0.5, 1.0,
                                     1.0 - 2**-53, 2.0, 4.0, 8.0, 0.125, 1e-300, 1e300,
                                     1.7976931348623157e308])])


def test_pow23_within_three_ulp(lib):
    """x**(2/3) by cube root (pow23, the canopy's power) against mpmath; exact
    cubes come out exact."""
    rng = np.random.default_rng(6)
    x = _pow23_sample(rng, 3000)
    x = x[x > 0]
    got = _eval(lib, 6, x)
    worst = max(_ulp_err(g, mp.power(mp.mpf(v), mp.mpf(2) / 3)) for g, v in zip(got, x))
    assert worst <= 3.0, worst
    cubes = np.array([2.0 ** (3 * k) for k in range(-340, 341)] + [27.0, 1 / 27.0, 125.0, 0.001])
    want = np.array([2.0 ** (2 * k) for k in range(-340, 341)] + [9.0, 1 / 9.0, 25.0, 0.01])
    got = _eval(lib, 6, cubes)
    assert np.array_equal(got[:681], want[:681])
    assert got[681] == 9.0 and got[683] == 25.0
    assert max(_ulp_err(g, mp.power(mp.mpf(v), mp.mpf(2) / 3)) for g, v in zip(got, cubes)) <= 3.0


@pytest.mark.gpu
def test_device_pow23_within_three_ulp(lib):
    """The device's pow23 seeds its cube root from the hardware's
    single-precision log2 / exp2: held to the same bound as the host."""
    print("case: pow23 on the device", flush=True)
    rng = np.random.default_rng(7)
    x = _pow23_sample(rng, 20000)
    x = x[x > 0]
    got = _eval(lib, 6, x, device=True)
    sub = np.arange(0, x.shape[0], 7)
    worst = max(_ulp_err(got[i], mp.power(mp.mpf(x[i]), mp.mpf(2) / 3)) for i in sub)
    assert worst <= 3.0, worst


@pytest.mark.gpu
def test_device_gives_host_bits(lib):
    print("case: pow device vs host bits", flush=True)
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(1e-9, 1.0, 50000), np.exp(rng.uniform(-700, 700, 50000))])
    y = rng.uniform(0.1, 10.0, x.shape[0])
    z = rng.uniform(-745, 709, x.shape[0])
    assert np.array_equal(_eval(lib, 0, x, device=True), _eval(lib, 0, x))
    assert np.array_equal(_eval(lib, 1, z, device=True), _eval(lib, 1, z))
    assert np.array_equal(_eval(lib, 2, x, y, device=True), _eval(lib, 2, x, y))
    assert np.array_equal(_eval(lib, 3, x, y, device=True), _eval(lib, 2, x, y))
    d = rng.uniform(0.3, 4000.0, x.shape[0])
    assert np.array_equal(_eval(lib, 4, x, d, device=True), x / d)

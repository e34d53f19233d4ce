"""Division by a model constant on the device (grw_model.h rdiv(double, DivK)):
q0 = RN(a rb), r = fma(-b, q0, a), q = fma(r, rb, q0) with rb = RN(1/b) must be
IEEE a / b bit for bit.  Markstein's theorem gives it for normal operands; here
the same three operations (Python float multiply, the C library's correctly
rounded fma) are checked against IEEE division for every divisor the device
code uses, on random dividends over a wide exponent range and on dividends
whose exact quotient lies within a few ulp^2 of a rounding midpoint, where a
wrong rounding would show."""
import ctypes
import ctypes.util
import random
import zlib
from fractions import Fraction

import numpy as np
import pytest

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.fma.restype = ctypes.c_double
_libm.fma.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
fma = _libm.fma

# the divisors of rdiv(x, divk(...)) in grw_fru.hip, grw_model.h and grw_global.h
DIVISORS = {
    "RT_over_F": 8.314 * 310.0 / 96.5,
    "tautr": 3.0,
    "tauxfer": 0.005,
    "tauss2ss": 10.0 * 0.005,
    "5": 5.0,
    "12": 12.0,
    "bL": 1.9356 * 1.0,
    "k_rate": 0.00127215,
    "k_rate2": 3.4188,
    # fcn_prep (grw_global.h)
    "Kfb": 0.26e-3,
    "Krb": 1.8,
    "5.98": 5.98,
    "11": 11.0,
    "-6.8": -6.8,
    "-11.1": -11.1,
    "13.6": 13.6,
    # volumes and the membrane area (fcn_assemble, flux_currents_e), the stage bounds
    "Vmyo": 25.84e-6,
    "VNSR": 0.53 * 2.10e-6,
    "Acap": 1.534e-4,
    "-2 Vmyo F 1000": -2.0 * 25.84e-6 * 96.5 * 1000.0,
    "10": 10.0,
    # the Ito1 lattice constants (grw_global.h fcn_ito_lattice)
    "f43_1": 1.8936, "f43_2": 14.224647456, "f43_3": 158.574378389, "f43_4": 142.936645351,
    "b43_1": 6.77348, "b43_2": 15.6212705152, "b43_3": 28.7532603313, "b43_4": 524.576206679,
    "f14_1": 0.20005, "f14_2": 0.320280, "f14_3": 13.50909223, "f14_4": 1151.7651385,
    "b14_1": 2.2300, "b14_2": 12.000299, "b14_3": 5.3701338025, "b14_4": 5.2396395511,
}


def divk(a: float, b: float) -> float:
    rb = 1.0 / b
    q0 = a * rb
    return fma(fma(-b, q0, a), rb, q0)


def test_formula_catches_a_plain_reciprocal_product():
    """The check has teeth: a * RN(1/b) alone is not IEEE division."""
    rng = random.Random(1)
    b = DIVISORS["tautr"]
    xs = [rng.uniform(-1.0, 1.0) for _ in range(2000)]
    assert sum(x * (1.0 / b) != x / b for x in xs) > 100
    assert all(divk(x, b) == x / b for x in xs)


@pytest.mark.parametrize("name", sorted(DIVISORS))
def test_constant_division_is_ieee(name):
    b = DIVISORS[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    # random dividends: mantissas in [1, 2), exponents over +-60 and both signs
    m = rng.uniform(1.0, 2.0, 20000)
    e = rng.integers(-60, 21, 20000)
    s = rng.choice([-1.0, 1.0], 20000)
    vals = list(s * np.ldexp(m, e)) + [0.0, -0.0, b, -b, 2 * b, b / 3]
    # dividends whose quotient lies next to a midpoint of two doubles
    r = random.Random(7)
    for _ in range(6000):
        q = float(np.ldexp(r.uniform(1.0, 2.0), r.randint(-40, 10)))
        ulp = np.spacing(q)
        mid = Fraction(q) + Fraction(ulp) / 2
        a = float(Fraction(b) * mid)  # nearest double to b * mid
        vals.append(a)
        vals.append(float(np.nextafter(a, np.inf)))
        vals.append(float(np.nextafter(a, -np.inf)))
    mism = [a for a in vals if divk(a, b) != a / b]
    assert not mism, f"{name}: {len(mism)} mismatches, e.g. {mism[:3]}"

"""The series quotients of the device IK01A (K0, K1 of the finite-depth
modes, FinGrnExtSubs.f90:538-628) are formed without FP64 division: for the
integer divisors b = k k and k (k + 1), k = 1..50, and rb = RN(1/b),

    q = RN(a rb);  r = a - q b  (exact, one FMA);  q' = RN(q + r rb)  (one FMA)

(Markstein's correction step).  The kernels rely on q' being the IEEE
quotient RN(a / b) for every argument, which keeps the assembled matrices bit
for bit those of the plain division.  Checked here in exact rational
arithmetic: the FMAs are evaluated exactly and rounded once, as the hardware
does, and compared with Python's correctly rounded division."""
from fractions import Fraction

import numpy as np


def fma(x, y, z):
    return float(Fraction(x) * Fraction(y) + Fraction(z))  # exact, then one rounding


def div_by_rcp(a, b, rb):
    q = a * rb
    r = fma(-q, b, a)
    assert Fraction(r) == Fraction(a) - Fraction(q) * Fraction(b)  # the remainder is exact
    return fma(r, rb, q)


def test_quotients_equal_ieee_division():
    rng = np.random.default_rng(5)
    divisors = sorted({k * k for k in range(1, 51)} | {k * (k + 1) for k in range(1, 51)})
    for b in divisors:
        b = float(b)
        rb = 1.0 / b
        mant = rng.integers(0, 1 << 52, size=300, dtype=np.int64)
        expo = rng.integers(-60, 60, size=300)
        args = [float(np.ldexp(1.0 + int(m) / 2.0 ** 52, int(e))) for m, e in zip(mant, expo)]
        # mantissas next to the divisor's multiples, where a wrong rounding shows first
        args += [b * (1.0 + j * 2.0 ** -52) for j in range(-8, 9)] + [b * 3.0 + j * 2.0 ** -40 for j in range(-8, 9)]
        for a in args:
            assert div_by_rcp(a, b, rb) == a / b, (a, b)


def test_a_plain_reciprocal_product_is_not_enough():
    """Without the correction step the product alone misses the IEEE quotient
    for some arguments, so the check above has something to catch."""
    rng = np.random.default_rng(6)
    b = 49.0
    rb = 1.0 / b
    args = rng.uniform(1.0, 2.0, size=2000)
    assert any(float(a) * rb != float(a) / b for a in args)

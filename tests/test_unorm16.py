"""The device's R16_UNORM decode (tc_device.h unorm16) replaces x / 65535.0f
by q0 = x * R, r = fma(-q0, 65535, x), q = fma(r, R, q0) with R the rounded
reciprocal.  Every one of the 65536 inputs is checked here against the IEEE
quotient, with each rounding done exactly (rational arithmetic); the GPU
parity tests then compare the displaced vertices that use it bit for bit."""
from fractions import Fraction

import numpy as np

F32 = np.float32


def round32(fr: Fraction) -> np.float32:
    """Nearest float32 to an exact rational (ties to even)."""
    f = F32(float(fr))
    cands = [np.nextafter(f, F32(-np.inf)), f, np.nextafter(f, F32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - fr), int(np.asarray(c, F32).view(np.uint32)) & 1))


def test_unorm16_matches_ieee_division_for_every_input():
    R = F32(1.0) / F32(65535.0)
    fr = Fraction(float(R))
    for x in range(65536):
        q0 = round32(Fraction(x) * fr)
        r = round32(Fraction(x) - Fraction(float(q0)) * 65535)
        q = round32(Fraction(float(r)) * fr + Fraction(float(q0)))
        assert q == F32(x) / F32(65535.0), x


def test_round32_is_exact_rounding():
    # a plain division through float64 double-rounds some quotients; the helper must not
    assert round32(Fraction(1, 3)) == F32(1.0) / F32(3.0)
    assert round32(Fraction(float(F32(1.0)) + float(np.finfo(np.float32).eps) / 2)) == F32(1.0)  # tie to even


def test_uncorrected_product_is_not_enough():
    # the residual step matters: x * R alone misses the quotient for some inputs
    x = np.arange(65536, dtype=F32)
    R = F32(1.0) / F32(65535.0)
    assert ((x * R).astype(F32) != x / F32(65535.0)).sum() > 0

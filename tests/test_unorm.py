"""The kernels' unorm8 -> float conversion (xd_device.hpp unorm8_to_float):
fma(c, h, fl(c * l)) with the two constants written in the header must be
c / 255 correctly rounded -- what the shaders' unpackUnorm4x8 / RGBA8_UNORM
fetch and the oracle's IEEE division give -- for every byte value.  Checked
here in exact rational arithmetic, with the constants read from the header."""
import os
import re
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _round32(x):
    """x (a Fraction) rounded to the nearest float32, ties to even."""
    f = np.float32(float(x))  # within one ulp of the answer
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda v: (abs(Fraction(float(v)) - x), int(v.view(np.uint32)) & 1))


def _header_constants():
    text = open(os.path.join(ROOT, "xenodon_amd", "csrc", "xd_device.hpp")).read()
    hi = re.search(r"kInv255Hi\s*=\s*([-0-9a-fx.p]+)f;", text).group(1)
    lo = re.search(r"kInv255Lo\s*=\s*([-0-9a-fx.p]+)f;", text).group(1)
    return np.float32(float.fromhex(hi)), np.float32(float.fromhex(lo))


def test_constants_are_the_split_of_one_over_255():
    h, lo = _header_constants()
    assert h == np.float32(1.0) / np.float32(255.0)
    assert lo == _round32(Fraction(1, 255) - Fraction(float(h)))


def test_conversion_is_correctly_rounded_for_every_byte():
    h, lo = _header_constants()
    for c in range(256):
        x = np.float32(c)
        cl = np.float32(x * lo)  # one float32 product
        got = _round32(Fraction(c) * Fraction(float(h)) + Fraction(float(cl)))  # the fma: one rounding
        assert got == _round32(Fraction(c, 255)), c
        assert got == np.float32(c) / np.float32(255), c  # IEEE division, as the oracle does


def test_a_plain_product_would_not_do():
    """The residual term matters: fl(c * h) alone misses half the bytes."""
    h, _ = _header_constants()
    wrong = sum(np.float32(c) * h != _round32(Fraction(c, 255)) for c in range(256))
    assert wrong > 100

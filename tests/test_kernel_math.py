"""CPU model of the fp64 engine's arithmetic (clb_device.hpp / clb_kernels.hip), with exactly
rounded fma, checked against exact integers.  It pins the error analysis the kernels rely on:
the balanced modular square stays exact and within 0.75 m, the lazy 48-bit quotient digits
reconstruct floor(r 2^192 / m), and the tail's long division of 2^E (whose first digit may
estimate -1) gives floor(2^E / m).  The GPU tests check the kernels themselves."""
import math
import random
from fractions import Fraction

import pytest

M192 = (1 << 192) - 1


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # one rounding, to nearest even


def recip(m: float, ulps: int) -> float:
    r = 1.0 / m
    for _ in range(abs(ulps)):
        r = math.nextafter(r, math.inf if ulps > 0 else -math.inf)
    return r


def f2u64(f: float) -> int:
    h = math.floor(f * 2.0**-32)
    lo = fma(h, -(2.0**32), f)
    return (((int(h) & 0xFFFFFFFF) << 32) | (int(lo) & 0xFFFFFFFF)) & ((1 << 64) - 1)


def fsq_shift(x: float, c: int, m: float, minv: float) -> float:
    y = math.ldexp(x, c)
    hi = x * y
    lo = fma(x, y, -hi)
    q = float(round(hi * minv))  # v_rndne_f64: round half to even, as Python's round
    t = fma(-q, m, hi)
    return t + lo


def fdigit(rem: float, m: float, minv: float):
    n = math.ldexp(rem, 48)
    q = math.floor(fma(n, minv, -0.25))
    return float(q), fma(-q, m, n)


def head_term(r: int, m: int, ulps: int) -> int:
    """Quotient digits of the head terms: floor(r 2^192 / m) from a balanced residue r."""
    mf, minv = float(m), recip(float(m), ulps)
    rem = float(r - m if r > m // 2 else r)
    total = 0
    for i in range(3, -1, -1):
        q, rem = fdigit(rem, mf, minv)
        total += int(q) << (48 * i)
    if rem >= mf:
        total += 1
    return total & M192


def tail_quot(E: int, m: int, ulps: int) -> int:
    """quot_pow2_f of clb_kernels.hip: floor(2^E / m) for m < 2^48, 2 <= E <= 192."""
    mf, minv = float(m), recip(float(m), ulps)
    a = 48 if E % 48 == 0 else E % 48
    c = (E - a) // 48
    rem, q = 1.0, 0
    for j in range(c + 1):
        n = math.ldexp(rem, a if j == 0 else 48)
        qd = max(float(math.floor(fma(n, minv, -0.25))), 0.0)
        rem = fma(-qd, mf, n)
        assert 0.0 <= rem < 2 * mf
        q += f2u64(qd) << (48 * (c - j))
    if rem >= mf:
        q += 1
    return q & M192


@pytest.mark.parametrize("W,bits", [(3, 42), (2, 46), (1, 48)])
def test_balanced_square_is_exact(W, bits):
    rng = random.Random(W)
    for _ in range(3000):
        m = rng.randrange(1 << (bits - 8), 1 << bits) | 1
        minv = recip(float(m), rng.choice((-1, 0, 1)))
        x = rng.randrange(-(3 * m) // 4, (3 * m) // 4 + 1)
        c = rng.randrange(0, (1 << W) + 1)  # window digit, plus the absorbed carry bit
        got = fsq_shift(float(x), c, float(m), minv)
        assert abs(got) <= 0.75 * m
        assert int(got) % m == (x * x << c) % m


def test_quotient_digits_reconstruct_the_term():
    rng = random.Random(11)
    for _ in range(3000):
        m = rng.randrange(3, 1 << 48) | 1
        r = rng.randrange(1, m)
        assert head_term(r, m, rng.choice((-1, 0, 1))) == (r << 192) // m


def test_tail_long_division():
    rng = random.Random(5)
    cases = [(E, m) for E in range(2, 193, 10) for m in (5, 9, 41, 203, 1 << 20 | 1)]
    cases += [(192 - 10 * rng.randrange(20), rng.randrange(5, 1 << 48) | 1) for _ in range(3000)]
    for E, m in cases:
        for ulps in (-1, 0, 1):
            assert tail_quot(E, m, ulps) == (1 << E) // m, (E, m, ulps)


def test_tail_first_digit_estimate_can_be_negative():
    """The case the clamp exists for: 2^a < m / 4 makes the biased estimate -1."""
    E, m = 52, 203  # a = 4
    n = math.ldexp(1.0, E % 48)
    assert math.floor(fma(n, recip(float(m), 0), -0.25)) == -1
    assert tail_quot(E, m, 0) == (1 << E) // m

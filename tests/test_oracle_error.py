"""Pins the 256-bit error oracle (orc_term256, orc_tail256, orc_job256, orc_normal256), against
which the fraction error of every 192-bit result is measured.

Checks: Python's exact integers (pow, //) term by term; the 192-bit oracle's terms are its top
192 bits (floor(floor(x 2^256) / 2^64) = floor(x 2^192)); the tail against the exact rational
series; BatchCalc's range and tail rule and the c_j factor against Python sums; and whole
positions against 256 bits of pi from mpmath (tests/golden/pi_hex.json, "digits256") at every
fixture position, where the oracle may lie below pi by at most its floor losses.
"""
import random
from fractions import Fraction

import pytest

from clb_testlib import (DEN0, DEN1, M256, NUMER, NUMESIGN, bound_ulps, frac_error_256,
                         py_normalize, py_tail256, py_term256)

R = random.Random(20261204)


def test_term256_exact_and_extends_term192(oracle):
    for _ in range(1500):
        j = R.randrange(7)
        d = R.choice([R.randrange(1, 10**4), R.randrange(1, 10**12), R.randrange(1, 10**17)])
        k = R.randrange(0, d)
        t = oracle.term256(d, k, NUMESIGN[j], DEN0[j], DEN1[j])
        assert t == py_term256(d, k, NUMESIGN[j], DEN0[j], DEN1[j]), (d, k, j)
        assert t >> 64 == oracle.term192(d, k, NUMESIGN[j], DEN0[j], DEN1[j]), (d, k, j)


def test_term256_m_equals_one_is_zero(oracle):
    for j in (0, 2):
        assert oracle.term256(11, 0, NUMESIGN[j], DEN0[j], DEN1[j]) == 0


def test_tail256_exact_and_against_the_series(oracle):
    for _ in range(200):
        j = R.randrange(7)
        d = R.randrange(1, 10**17)
        assert oracle.tail256(d, NUMESIGN[j], DEN0[j], DEN1[j]) == py_tail256(d, j)
    # the tail to k = d + 25 is the whole (infinite) tail of the series to within 27 units
    for d, j in ((12345, 3), (99, 0), (399999999999, 6)):
        exact = sum(Fraction((-1) ** (d + i + NUMESIGN[j]),
                             (2 ** (10 * i)) * (DEN0[j] * (d + i) + DEN1[j])) for i in range(60))
        got = oracle.tail256(d, NUMESIGN[j], DEN0[j], DEN1[j])
        diff = (exact * 2**256 - got) % 2**256
        assert diff < 27, (d, j, float(diff))


def test_job256_rule_and_factor(oracle):
    d = 301
    for j in range(7):
        for k0, k1 in [(0, d), (0, 10**6), (100, d), (100, 200), (d, d + 5), (d + 1, d + 9),
                       (0, 0), (250, 300)]:
            hi = min(k1, d)
            s, offset = 0, k0
            if k0 < hi:
                s = sum(py_term256(d, k, NUMESIGN[j], DEN0[j], DEN1[j]) for k in range(k0, hi))
                offset = hi
            if offset == d:
                s += py_tail256(d, j)
            assert oracle.job256(d, k0, k1, j) == (s * NUMER[j]) & M256, (j, k0, k1)
    # partition invariance and thread invariance
    d, j = 100001, 4
    cuts = [0] + sorted(R.sample(range(1, d), 5)) + [d]
    total = sum(oracle.job256(d, a, b, j, threads=1) for a, b in zip(cuts, cuts[1:])) & M256
    assert total == oracle.job256(d, 0, d, j, threads=7)


POSITIONS = [5, 10, 15, 20, 25, 30, 50, 100, 260, 500, 1000, 1005, 4100, 9995, 10000, 65540,
             100000, 123455, 250000, 500000, 777775, 1000000]


@pytest.mark.parametrize("n", POSITIONS)
def test_normal256_is_pi(oracle, pi_fixtures, n):
    """The oracle's 256-bit value lies below pi's 256 bits by at most its floor losses, one unit
    per term times c_j (362 (d + 26) units of 2^-256); and the 192-bit result's measured error
    against it is within the 192-bit bound."""
    _, d = py_normalize(n)
    pi256 = int(pi_fixtures["digits256"][str(n)], 16)
    got = oracle.normal256(n)
    assert 0 <= (pi256 - got) % (1 << 256) <= sum(NUMER) * (d + 26) + 2, n
    r192 = oracle.normal(n)
    err = frac_error_256(r192, pi256)
    assert -(bound_ulps(d) << 64) <= err <= 0, (n, err)
    assert abs(frac_error_256(r192, got) - err) <= sum(NUMER) * (d + 26) + 2

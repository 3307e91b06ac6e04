"""Pins the CPU oracle (oracle/clb_oracle.c) before the parity tests trust it.

Each oracle function is checked against something the mathematics fixes: closed forms, Python's
exact integers (pow, //), Bellard's series for pi as an exact rational, and the hex digits of pi
from mpmath (tests/golden/pi_hex.json).  The reference has no tests or vectors of its own.
"""
import ctypes
import math
import random
from fractions import Fraction

import pytest

from clb_testlib import (DEN0, DEN1, M192, NUMER, NUMESIGN, U3, U4, bound_ulps, i2w,
                         pi_series_check, py_batch_calc, py_normal, py_normalize, py_range_sum,
                         py_tail, py_tail_term, py_term, w2i)

R = random.Random(20261203)


# ------------------------------------------------------------------ host constants

def test_normalize_closed_form(oracle):
    # Bellard_192.hsp:468-470: d = ((N-1)*2-3)/5 forced odd, N' = (5d+3)/2+1, 10d+6 = 4(N'-1)
    for n, (nn, d) in {5: (5, 1), 6: (5, 1), 9: (5, 1), 10: (10, 3), 11: (10, 3), 1000: (1000, 399),
                       250000: (250000, 99999), 10**12: (10**12, 399999999999),
                       2500000000000: (2500000000000, 999999999999)}.items():
        got = oracle.normalize(n)
        assert got[:2] == (nn, d), n
        assert 10 * d + 6 == 4 * (nn - 1) and nn % 5 == 0 and d % 2 == 1
    for _ in range(2000):
        n = R.randrange(5, 10**18)
        nn, d, od = oracle.normalize(n)
        assert (nn, d) == py_normalize(n)
        assert od == (10 * d + 9 >= 2**53)
        assert nn <= n and n - nn < 5 + 5  # rounds down to a multiple of 5 (within one step)
    assert oracle.normalize(4) is None and oracle.normalize(10**18) is None
    assert oracle.normalize(999999999999999999) is not None


def test_formula_table_is_bellards_series(oracle):
    # The seven (c_j, sign_j, den0_j, den1_j) must sum to pi: 12 terms give ~120 bits.
    got = []
    for j in range(7):
        c, s, a, b = ctypes.c_uint64(), ctypes.c_int64(), ctypes.c_uint64(), ctypes.c_uint64()
        assert oracle.lib.orc_formula(j, ctypes.byref(c), ctypes.byref(s), ctypes.byref(a),
                                      ctypes.byref(b)) == 0
        got.append((c.value, s.value, a.value, b.value))
    assert got == list(zip(NUMER, NUMESIGN, DEN0, DEN1))
    approx = pi_series_check(12)
    pi_ref = Fraction(int("3243f6a8885a308d313198a2e03707344a4093822299f31d0082efa98ec4e6c8", 16),
                      1 << 252)
    assert abs(approx - pi_ref) < Fraction(1, 1 << 110)


# ------------------------------------------------------------------ Montgomery arithmetic

def test_modinv(oracle):
    for _ in range(2000):
        m32 = R.randrange(1, 2**32, 2)
        assert (m32 * oracle.lib.orc_modinv32(m32)) % 2**32 == 2**32 - 1
        m64 = R.randrange(1, 2**64, 2)
        assert (m64 * oracle.lib.orc_modinv64(m64)) % 2**64 == 2**64 - 1


def test_mr64_is_redc(oracle):
    for _ in range(2000):
        m = R.randrange(3, 2**63, 2)
        x = R.randrange(0, m * 2**64)
        inv = oracle.lib.orc_modinv64(m)
        got = oracle.lib.orc_mr64(x % 2**64, x >> 64, inv, m)
        assert got == (x * pow(2, -64, m)) % m


def test_expmod32_brute_force(oracle):
    for m in list(range(1, 200, 2)) + [R.randrange(1, 2**32, 2) for _ in range(300)]:
        inv = oracle.lib.orc_modinv32(m)
        r2 = (2**64 - 1) % m + 1
        if r2 == m:
            r2 = 0
        for n in (0, 1, 2, 3, 7, R.randrange(0, 2**34)):
            assert oracle.lib.orc_expmod32(n, m, inv, r2) == pow(1024, n, m), (m, n)


def test_expmod64_brute_force(oracle):
    for _ in range(400):
        m = R.randrange(2**32 + 1, 2**62, 2)
        inv = oracle.lib.orc_modinv64(m)
        r2 = pow(2, 128, m)
        for n in (0, 1, 2, 5, 8, 9, R.randrange(0, 2**45)):
            assert oracle.lib.orc_expmod64(n, m, inv, r2) == pow(1024, n, m), (m, n)


def test_pow1024_mod_both_paths(oracle):
    ms = [1, 3, 5, 2**32 - 1, 2**32 + 1, 2**61 - 1] + [R.randrange(1, 2**62, 2) for _ in range(500)]
    for m in ms:
        n = R.randrange(0, 2**42)
        assert oracle.pow1024_mod(n, m) == pow(1024, n, m), m


# ------------------------------------------------------------------ division

def test_reciprocal_word(oracle):
    ds = [2**63, 2**63 + 1, 2**64 - 1] + [R.randrange(2**63, 2**64) for _ in range(3000)]
    for d in ds:
        assert oracle.lib.orc_reciprocal_word(d) == (2**128 - 1) // d - 2**64, d


def test_div2by1(oracle):
    for _ in range(3000):
        d = R.randrange(2**63, 2**64)
        u1 = R.randrange(0, d)
        u0 = R.randrange(0, 2**64)
        v = oracle.lib.orc_reciprocal_word(d)
        r = ctypes.c_uint64()
        q = oracle.lib.orc_div2by1(u1, u0, d, ctypes.byref(r), v)
        assert (q, r.value) == divmod((u1 << 64) | u0, d)


def test_divnby1_and_mulmodn(oracle):
    for _ in range(1500):
        d = R.randrange(1, 2**64)
        u = R.randrange(0, 2**256)
        uw = U4(*[(u >> (64 * i)) & (2**64 - 1) for i in range(4)])
        qw = U4()
        oracle.lib.orc_divnby1(4, uw, d, qw)
        assert sum(int(qw[i]) << (64 * i) for i in range(4)) == u // d
        n = R.randrange(2, 2**64)
        s = 64 - n.bit_length()
        v = oracle.lib.orc_reciprocal_word(n << s)
        x, y = R.randrange(0, n), R.randrange(0, n)
        assert oracle.lib.orc_mulmodn(x, y, n, s, v) == (x * y) % n


# ------------------------------------------------------------------ terms and sums

def test_term192_exact(oracle):
    for _ in range(1500):
        j = R.randrange(7)
        d = R.choice([R.randrange(1, 10**4), R.randrange(1, 10**12), R.randrange(1, 10**17)])
        k = R.randrange(0, d)
        assert oracle.term192(d, k, NUMESIGN[j], DEN0[j], DEN1[j]) == \
            py_term(d, k, NUMESIGN[j], DEN0[j], DEN1[j]), (d, k, j)


def test_term192_m_equals_one_is_zero(oracle):
    # k = 0 of 4k+1 and 10k+1: m = 1, r = 0; the negative encoding m - r = 1 gives 2^192 = 0.
    for j in (0, 2):
        assert oracle.term192(11, 0, NUMESIGN[j], DEN0[j], DEN1[j]) == 0


def test_tail_exact(oracle):
    for _ in range(300):
        j = R.randrange(7)
        d = R.randrange(1, 10**17)
        assert oracle.tail192(d, NUMESIGN[j], DEN0[j], DEN1[j]) == py_tail(d, j)
    # the tail is the rational sum of the 20 terms to within 20 units
    d, j = 12345, 3
    exact = sum(Fraction((-1) ** (d + i + NUMESIGN[j]), (2 ** (10 * i)) * (DEN0[j] * (d + i) + DEN1[j]))
                for i in range(20))
    got = oracle.tail192(d, NUMESIGN[j], DEN0[j], DEN1[j])
    diff = (exact * 2**192 - got) % 2**192
    assert diff < 20 or diff > 2**192 - 1


def test_negative_term_is_complement(oracle):
    # frac(-x) = 1 - x: the reference's m - r encoding is the 192-bit complement of the floor
    d, k, den0, den1 = 999, 10, 10, 3
    pos = oracle.term192(d, k, 0, den0, den1) if k % 2 == 0 else oracle.term192(d, k, 1, den0, den1)
    neg = oracle.term192(d, k, 1, den0, den1) if k % 2 == 0 else oracle.term192(d, k, 0, den0, den1)
    assert (pos + neg) & M192 == M192


def test_range_sum_brute_force_and_partition_invariance(oracle):
    for _ in range(30):
        j = R.randrange(7)
        d = R.randrange(1, 4000) | 1
        k0 = R.randrange(0, d)
        k1 = R.randrange(k0, d + 1)
        assert oracle.range_sum(d, k0, k1, j) == py_range_sum(d, k0, k1, j)
    d, j = 100001, 5
    whole = oracle.range_sum(d, 0, d, j)
    cuts = sorted(R.sample(range(1, d), 7))
    parts = [0] + cuts + [d]
    total = sum(oracle.range_sum(d, a, b, j, threads=1) for a, b in zip(parts, parts[1:]))
    assert total & M192 == whole
    assert oracle.range_sum(d, 0, d, j, threads=1) == oracle.range_sum(d, 0, d, j, threads=7)


def test_batch_calc_tail_rule(oracle):
    # KernelFuncs.as:56-59: the tail runs iff the final offset equals d
    d = 301
    for j in range(7):
        for k0, k1 in [(0, d), (0, 10**6), (100, d), (100, 200), (d, d + 5), (d, 0), (d + 1, d + 9),
                       (0, 0), (250, 300)]:
            assert oracle.batch_calc(d, k0, k1, j) == py_batch_calc(d, k0, k1, j), (j, k0, k1)


def test_sum7_shift(oracle):
    for _ in range(500):
        ans, part = R.randrange(0, 2**192), R.randrange(0, 2**192)
        c = R.choice(NUMER)
        new_ans, new_part = oracle.sum7(ans, part, c)
        assert new_part == (part * c) & M192
        assert new_ans == (ans + part * c) & M192


# ------------------------------------------------------------------ whole pipeline vs pi

@pytest.mark.parametrize("n", [5, 10, 15, 20, 25, 30, 50, 100, 260, 500, 1000, 1005, 4100, 9995,
                               10000, 65540, 100000, 123455, 250000])
def test_normal_matches_pi_digits(oracle, pi_fixtures, n):
    _, d = py_normalize(n)
    got = oracle.normal(n)
    ref = int(pi_fixtures["digits"][str(n)], 16)
    # exact-integer semantics underestimate by the per-term floors only
    assert (ref - got) % (1 << 192) <= bound_ulps(d), (n, hex(got), hex(ref))


def test_normal_matches_python_big_integers(oracle):
    for n in (5, 10, 55, 100, 335):
        assert oracle.normal(n) == py_normal(n)


# ------------------------------------------------------------------ formats

def test_hex48(oracle):
    for _ in range(200):
        v = R.randrange(0, 2**192)
        assert oracle.hex48(v) == f"{v:048x}"


def _dd_exact_hex(hi, lo):
    v = Fraction(hi) + Fraction(lo)
    if v < 0:
        v += 1
    top = int(v * 2**112)
    return f"{top:028x}" + "0" * 20


def test_hex_dd_exact_inputs(oracle):
    # Proper double-doubles (|lo| <= ulp(hi)/2) whose value has no bits below 2^-112 print
    # exactly.  With an odd mantissa in hi, a negative lo merges into a non-zero h (TwoSum)
    # before hi's bits run out, so the borrow path of the reference is never entered.
    cases = [(0.5, 0.0), (0.75, 2.0**-100), (-0.25, 0.0), (0.3, 0.0), (0.99609375, 2.0**-60),
             (1.0 - 2.0**-53, -2.0**-57), (-0.5, 2.0**-70)]
    for _ in range(500):
        hi = (R.randrange(2**52, 2**53) | 1) / 2.0**53
        lo = R.choice([1, -1]) * R.randrange(2**44, 2**50) / 2.0**107
        cases.append((hi, lo))
        cases.append((hi, abs(lo) / 2.0**7))
    for hi, lo in cases:
        assert oracle.hex_dd(hi, lo) == _dd_exact_hex(hi, lo), (hi, lo)


def test_hex_dd_general_inputs(oracle):
    """Double-doubles with bits far below 2^-112.  hexsademical.as:38-71 keeps hh + ll exact
    (x256 is exact, TwoSum is error-free, hh - digit is exact), so a non-negative value prints
    as floor(v 2^112) unless some step's remainder lies within half an ulp below an integer (46
    one-bits in a row; random mantissas never do).  A negative value (a0, a1) first gets 1 added
    by addcab (:12-19, :25-31): TwoSum(a0, 1) = (x, y) is exact, the one rounding is
    s = RN(y + a1) (the + 0.0 of the second operand's low part is exact), and the closing TwoSum
    is exact, so the value printed is floor((x + s) 2^112).  Computed here with Python's own
    IEEE round-to-nearest additions and exact rationals; the difference from the exact
    floor((v + 1) 2^112) is that one rounding, |s - (y + a1)| <= 2^-107 (up to 2^5 units)."""
    rng = random.Random(2024)
    moved = 0
    for _ in range(3000):
        hi = rng.randrange(2**52, 2**53) / 2.0**53 * 2.0 ** -rng.randrange(0, 20)
        lo = math.ldexp(rng.random() - 0.5, math.frexp(hi)[1] - 53)  # |lo| <= ulp(hi) / 2
        want = int((Fraction(hi) + Fraction(lo)) * 2**112)
        assert oracle.hex_dd(hi, lo) == f"{want:028x}" + "0" * 20, (hi, lo)
        a0, a1 = -hi, -lo
        x = a0 + 1.0
        y = float(Fraction(a0) + 1 - Fraction(x))  # TwoSum's error term: exact
        assert Fraction(y) == Fraction(a0) + 1 - Fraction(x)
        s = y + a1
        want_neg = math.floor((Fraction(x) + Fraction(s)) * 2**112)
        assert oracle.hex_dd(a0, a1) == f"{want_neg:028x}" + "0" * 20, (hi, lo)
        exact = math.floor((1 + Fraction(a0) + Fraction(a1)) * 2**112)
        assert abs(want_neg - exact) <= 33
        moved += want_neg != exact
    # the rounding of y + a1 reaches the printed bits in most cases: the rule above is pinned by
    # cases where it matters, not only by cases where it vanishes
    assert moved > 1000


def test_hex_dd_reference_borrow_quirk(oracle):
    # hexsademical.as:73-97: when the remaining value h goes negative the byte becomes
    # int(h + 256.0) and a borrow is taken, but h is only increased by 256 -- the emitted byte is
    # not subtracted -- so h stays in (255, 256] and no later byte is written.  Traced by hand:
    #   (0.5, -2^-60): byte 0x80, then h = -2^-44 -> byte 0xff with borrow -> "7fff" + zeros
    #   (0.5, -2^-80): byte 0x80, then h = -2^-64, 256 - 2^-64 rounds to 256 -> byte 0x00 with
    #                  borrow -> "7f" + zeros
    assert oracle.hex_dd(0.5, -2.0**-60) == "7fff" + "0" * 44
    assert oracle.hex_dd(0.5, -2.0**-80) == "7f" + "0" * 46
    assert oracle.hex_dd(-0.5, -2.0**-60) == "7fff" + "0" * 44  # -0.5 + 1 first (:25-31)


def test_work_sum_add(oracle):
    acc = (ctypes.c_int * 48)()
    total = 0
    for _ in range(40):
        v = R.randrange(0, 2**192)
        total += v
        oracle.lib.orc_work_sum_add(acc, f"{v:048x}".encode())
    got = int("".join("0123456789abcdef"[x] for x in acc), 16)
    assert got == total % 2**192

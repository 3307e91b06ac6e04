"""CPU: the work counts behind bench.py's roofline, against brute force over every term.

The roofline's `achieved` is an algorithmic count divided by the head kernel's time, so the counts
are pinned here term by term on small loop counts and at the bit-length boundaries of the
exponents 10 (d - k): the sum of exponent bit lengths (squares per term = bitlen - 1), the FP64
issue slots (6 per square, 16 per term's digits), the word products of a 2-limb Montgomery engine
(10 per square and per quotient word), and the plain kernel's multiply instructions (its
Montgomery exponent 10 (d - k) + 192 taken in 3-bit windows from the top)."""
import os
import random
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def brute_bitlen_sum(d):
    return sum((10 * x).bit_length() for x in range(1, d + 1))


@pytest.mark.parametrize("d", [1, 2, 3, 6, 7, 12, 13, 25, 26, 99, 1000, 4096, 65535, 99999])
def test_sum_bitlen_10x_small(d):
    assert bench.sum_bitlen_10x(d) == brute_bitlen_sum(d)


def test_sum_bitlen_10x_at_power_of_two_boundaries():
    # d around the points where 10 d crosses 2^b, where an off-by-one in the ranges would show
    for b in range(4, 18):
        for d in (((1 << b) - 1) // 10, ((1 << b) - 1) // 10 + 1):
            assert bench.sum_bitlen_10x(d) == brute_bitlen_sum(d), d


def test_sum_bitlen_10x_is_additive():
    # large d: differences of the closed form equal the bit lengths of the added terms
    rng = random.Random(7)
    for _ in range(50):
        d = rng.randrange(10**9, 10**12)
        step = rng.randrange(1, 200)
        added = sum((10 * x).bit_length() for x in range(d + 1, d + step + 1))
        assert bench.sum_bitlen_10x(d + step) - bench.sum_bitlen_10x(d) == added


def test_slot_and_product_counts_term_by_term():
    for d in (5, 40, 1000, 39999):
        squares = sum((10 * (d - k)).bit_length() - 1 for k in range(d))
        assert bench.algorithmic_fp64_slots(d) == 7 * sum(
            6 * ((10 * (d - k)).bit_length() - 1) + 16 for k in range(d))
        assert bench.algorithmic_fp64_flops(d) == 7 * (7 * squares + 20 * d)
        assert bench.int32_word_products(d) == 7 * 10 * (squares + 3 * d)


def test_plain_multiplies_term_by_term():
    rng = random.Random(3)
    cases = [(1000, 0, 1000), (39999, 0, 39999), (39999, 123, 4567)]
    for _ in range(20):
        d = rng.randrange(10**6, 10**12)
        k0 = rng.randrange(0, d - 3000)
        cases.append((d, k0, k0 + rng.randrange(1, 3000)))
    for d, k0, k1 in cases:
        want = 0
        for k in range(k0, k1):
            windows = ((10 * (d - k) + 192).bit_length() + 2) // 3
            want += bench.PLAIN_MULS_PER_WINDOW * (windows - 1) + bench.PLAIN_MULS_PER_TERM
        assert bench.plain_multiplies(d, k0, k1) == want, (d, k0, k1)


def test_headline_count():
    # the figure the roofline divides at N = 10^12 (d = 399999999999)
    assert bench.algorithmic_fp64_slots(399999999999) == 715128204652386

"""CPU model of the fp64 engine's arithmetic (clb_device.hpp / clb_kernels.hip), with exactly
rounded fma, checked against exact integers.  It pins the error analysis the kernels rely on:
the balanced modular squares (plain, windowed, doubled) stay exact and within 1.02-1.34 m, the
exponent chain gives 2^e mod m, the lazy 48-bit quotient digits and the per-lane digit sums
reconstruct floor(r 2^192 / m), the tail's long division of 2^E (whose first digit may
estimate -1) gives floor(2^E / m), the reference's fused Dekker product (which the replay
computes with one fma) is error-free, and the replay's signed numerator gives the reference's
negated quotient.  The GPU tests check the kernels themselves."""
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


def f2i64(f: float) -> int:
    """The kernel's conversion of a possibly negative digit sum: high half through a signed
    32-bit conversion, low half unsigned."""
    h = math.floor(f * 2.0**-32)
    lo = fma(h, -(2.0**32), f)
    assert -(2**31) <= h < 2**31 and 0 <= lo < 2**32
    return (((int(h) & 0xFFFFFFFF) << 32) | int(lo)) & ((1 << 64) - 1)


def fsq_shift(x: float, c: int, m: float, minv: float) -> float:
    y = math.ldexp(x, c)
    hi = x * y
    lo = fma(x, y, -hi)
    q = float(round(hi * minv))  # v_rndne_f64: round half to even, as Python's round
    t = fma(-q, m, hi)
    return t + lo


def fsq2(x: float, m: float, minv: float) -> float:
    """The doubling square of fsq_bits: fsq with the output modifier mul:2 on the final add."""
    hi = x * x
    lo = fma(x, x, -hi)
    q = float(round(hi * minv))
    t = fma(-q, m, hi)
    return 2.0 * (t + lo)


def fpow2_split(H: int, v: int, tmax: int, W: int, m: float, minv: float, ymax: float) -> float:
    """The kernel's 2^(1024 H + v) mod m: start 2^t, H's bits by plain or doubling squares, H's
    last bit with v's bit 10, then W-bit windows over v's low ten bits (v even, so for W = 3 and
    1 the last square takes bit 0 without a shift)."""
    bl, bt = H.bit_length(), tmax.bit_length()
    s = 0
    if bl > bt or (bl == bt and H >= tmax):
        s = bl - bt if bl > bt else 1
        if (H >> s) > tmax:
            s += 1
    t, carry = H >> s, v >> 10
    assert t + (carry if s == 0 else 0) <= tmax
    if s == 0:
        p = math.ldexp(1.0, t + carry)
        x = fma(-float(round(p * minv)), m, p)
    else:
        p = math.ldexp(1.0, t)
        x = fma(-float(round(p * minv)), m, p)
        for pos in range(s - 1, 0, -1):
            x = fsq2(x, m, minv) if (H >> pos) & 1 else fsq_shift(x, 0, m, minv)
            assert abs(x) <= ymax * m
        x = fsq_shift(x, (H & 1) + carry, m, minv)
    last = 0 if W == 2 else 1  # v is even: a window of bit 0 alone is a plain square
    for pos in range(10 - W, last - 1, -W):
        for _ in range(W - 1):
            x = fsq_shift(x, 0, m, minv)
        x = fsq_shift(x, (v >> pos) & ((1 << W) - 1), m, minv)
        assert abs(x) <= ymax * m
    if last == 1:
        assert v % 2 == 0
        x = fsq_shift(x, 0, m, minv)
    return x


def fdigit(rem: float, m: float, minv: float):
    n = math.ldexp(rem, 48)
    q = math.floor(fma(n, minv, -0.25))
    return float(q), fma(-q, m, n)


def head_term(x: int, m: int, ulps: int) -> int:
    """Quotient digits of the head terms: floor(r 2^192 / m) mod 2^192, r = x mod m, from the
    exponentiation's result x (|x| up to 1.34 m)."""
    mf, minv = float(m), recip(float(m), ulps)
    rem = float(x)
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


# (window width, modulus bits, bound on |x| / m kept by every step, bound after fsq_shift): with
# |x| <= y m the estimate of x y' / m is off by at most y^2 (m 2^c) 1.5 2^-52, which fixes y.
BOUNDS = [(3, 42, 1.02, 0.97), (2, 46, 1.06, 1.0), (1, 48, 1.34, 1.3)]


@pytest.mark.parametrize("W,bits,ymax,yshift", BOUNDS)
def test_balanced_square_is_exact(W, bits, ymax, yshift):
    """Inputs up to ymax m (what a doubling square can leave): fsq_shift stays exact and within
    yshift m, the doubling square within ymax m."""
    rng = random.Random(W)
    for trial in range(3000):
        m = rng.randrange(1 << (bits - 8), 1 << bits) | 1
        if trial % 10 == 0:
            m = (1 << bits) - 1 - 2 * rng.randrange(1 << 8)  # the largest moduli of the range
        minv = recip(float(m), rng.choice((-1, 0, 1)))
        x = rng.randrange(-int(ymax * m), int(ymax * m) + 1)
        if trial % 7 == 0:
            x = int(ymax * m) * rng.choice((-1, 1))
        c = rng.randrange(0, (1 << W) + 1)  # window digit, plus the absorbed carry bit
        got = fsq_shift(float(x), c, float(m), minv)
        assert abs(got) <= yshift * m
        assert int(got) % m == (x * x << c) % m
        got2 = fsq2(float(x), float(m), minv)
        assert abs(got2) <= ymax * m
        assert int(got2) % m == (2 * x * x) % m


@pytest.mark.parametrize("W,bits,ymax,yshift", BOUNDS)
def test_exponent_chain(W, bits, ymax, yshift):
    """The whole split-form exponentiation against pow(2, 1024 H + v, m), including exponents
    whose H is all ones (a doubling square on every bit)."""
    rng = random.Random(100 + W)
    for trial in range(400):
        m = rng.randrange(1 << (bits - 10), 1 << bits) | 1
        tmax = 49 + m.bit_length() - 1
        H = rng.randrange(1, 1 << rng.randrange(1, 40))
        if trial % 5 == 0:
            H = (1 << rng.randrange(1, 40)) - 1
        v = 2 * rng.randrange(5, 832)  # v = ((E - 640) & 1023) + 640 - 10 L, E even
        x = fpow2_split(H, v, tmax, W, float(m), recip(float(m), rng.choice((-1, 0, 1))), ymax)
        assert abs(x) <= ymax * m
        assert int(x) % m == pow(2, 1024 * H + v, m)


def test_start_choice_is_minimal():
    """s is the smallest shift with H >> s <= tmax (s = 0 only when H + 1 <= tmax, which leaves
    room for v's bit 10)."""
    for tmax in (49, 60, 87, 90, 96, 112):
        for H in list(range(1, 600)) + [(1 << b) + d for b in range(10, 40) for d in (-1, 0, 1)]:
            bl, bt = H.bit_length(), tmax.bit_length()
            s = 0
            if bl > bt or (bl == bt and H >= tmax):
                s = bl - bt if bl > bt else 1
                if (H >> s) > tmax:
                    s += 1
            want = 0 if H + 1 <= tmax else min(k for k in range(1, 64) if (H >> k) <= tmax)
            assert s == want, (H, tmax)


def tile_sum(xs, ms, ulps) -> int:
    """One lane's tile: four fp64 digit sums over up to 16 terms, the integer count of lazy last
    digits, converted (digit 0 as two's complement) and placed at 2^144, 2^96, 2^48, 2^0."""
    f = [0.0, 0.0, 0.0, 0.0]
    corr = 0
    for x, m in zip(xs, ms):
        mf, minv = float(m), recip(float(m), ulps)
        rem = float(x)
        for i in range(4):
            q, rem = fdigit(rem, mf, minv)
            f[i] += q
        corr += rem >= mf
    u3 = (f2u64(f[3]) + corr) & ((1 << 64) - 1)
    u0 = f2i64(f[0])
    return (u0 << 144) + (f2u64(f[1]) << 96) + (f2u64(f[2]) << 48) + u3 & M192


def test_tile_digit_sums():
    """16 terms per lane with residues anywhere in [-1.34 m, 1.34 m] (the widest the exponent
    chain leaves), all at the extremes too: the converted sums equal the exact floor sum."""
    rng = random.Random(12)
    for trial in range(300):
        ms = [rng.randrange(3, 1 << 48) | 1 for _ in range(16)]
        if trial % 3 == 0:
            xs = [int(1.34 * m) * (1 if trial % 2 else -1) for m in ms]
        else:
            xs = [rng.randrange(-int(1.34 * m), int(1.34 * m) + 1) for m in ms]
        want = sum(((x % m) << 192) // m for x, m in zip(xs, ms)) & M192
        assert tile_sum(xs, ms, rng.choice((-1, 0, 1))) == want, trial


def ulps_from_reciprocal(r: float, m: float) -> int:
    """Signed distance of r from the correctly rounded 1/m, in ulps (capped at 9)."""
    t = 1.0 / m
    n, x = 0, t
    while x != r and n < 9:
        x = math.nextafter(x, math.inf if r > t else -math.inf)
        n += 1
    return n if r > t else -n


def newton_update(m: float, r: float, steps: int) -> float:
    for _ in range(steps):
        e = fma(-m, r, 1.0)
        r = fma(r, e, r)
    return r


@pytest.mark.parametrize("steps,lo,hi", [(1, 37, 48), (2, 23, 37)])
def test_reciprocal_carried_along_the_stride(steps, lo, hi):
    """clb_kernels.hip split_terms: a lane's terms step m by 64 den0 (den0 4 or 10), and 1/m of the
    next term is taken from this term's reciprocal by one Newton step once m >= 2^37 and by two
    once m >= 2^23.  Over a tile's 16 terms, from a first reciprocal within an ulp of 1/m (the
    hardware estimate and two steps), every reciprocal stays within an ulp of the correctly
    rounded 1/m -- the accuracy the squares' and digits' bounds assume (recip(m, -1..1) above)."""
    rng = random.Random(steps)
    starts = [(1 << lo) + 1, (1 << lo) + 641, (1 << hi) - 64 * 10 * 16 - 1]
    starts += [rng.randrange(1 << lo, 1 << hi) | 1 for _ in range(600)]
    for m0 in starts:
        for den0 in (4, 10):
            r = recip(float(m0), rng.choice((-1, 0, 1)))
            m = m0
            for _ in range(15):
                m += 64 * den0
                r = newton_update(float(m), r, steps)
                assert abs(ulps_from_reciprocal(r, float(m))) <= 1, (m0, den0, m)


def test_one_newton_step_is_not_enough_below_2_37():
    """The bound that sets the switch: from 2^23 to 2^30 a single step leaves the reciprocal of the
    next modulus several ulps away (relative error (640 / m)^2 > 2^-53), so the tile choice of
    two steps there is needed, not a margin."""
    rng = random.Random(3)
    far = 0
    for _ in range(300):
        m0 = rng.randrange(1 << 23, 1 << 30) | 1
        r = 1.0 / m0
        m = m0 + 640
        far += abs(ulps_from_reciprocal(newton_update(float(m), r, 1), float(m))) > 1
    assert far > 250


def test_quotient_digits_reconstruct_the_term():
    rng = random.Random(11)
    for _ in range(3000):
        m = rng.randrange(3, 1 << 48) | 1
        x = rng.randrange(-int(1.34 * m), int(1.34 * m) + 1)
        assert head_term(x, m, rng.choice((-1, 0, 1))) == ((x % m) << 192) // m


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


def _dekker_fused(a: float, b: float):
    """twoproduct of the reference (Bellard.cl:17-27) as its build runs it: Veltkamp halves, the
    error term's four products fused in the order (ah bh - p) + al bh + ah bl + al bl."""
    def split(x):
        t = x * 134217729.0
        hi = t - (t - x)
        return hi, x - hi
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    e = fma(ah, bh, -p)
    e = fma(al, bh, e)
    e = fma(ah, bl, e)
    e = fma(al, bl, e)
    return p, e


def test_dekker_product_is_error_free():
    """The replay computes twoproduct as (a b, fma(a, b, -a b)): the same two doubles, because
    Dekker's product is error-free.  Checked on the operands the replay divides with: quotients
    x / m of either sign down to 2^-300 against integer moduli m < 2^53, and hard cases."""
    rng = random.Random(17)
    cases = [(0.5, 3.0), (-(2.0 ** -300), 9.0), (1.0 - 2.0**-53, 2.0**53 - 1), (2.0**-53, 2.0**52 + 1)]
    for _ in range(20000):
        m = rng.randrange(1, 1 << rng.choice((20, 33, 42, 48, 53)))
        a = rng.random() * 2.0 ** -rng.randrange(0, 300) * rng.choice((1, -1))
        cases.append((a, float(m)))
    for a, b in cases:
        p, e = _dekker_fused(a, b)
        assert p == a * b
        assert Fraction(e) == Fraction(a) * Fraction(b) - Fraction(p), (a, b)
        assert e == fma(a, b, -p)


def test_nn_div_final_sum_is_fast():
    """nn_div's last step two_sum(z1, z2) may be Fast2Sum: with z1 = RN(x / y) (correctly rounded,
    as the device division is) and z2 the rounded quotient of the residual, |z2| <= |z1|, and
    both sums return the same two doubles.  Numerators in [0, y], integer moduli y < 2^53."""
    def two_sum(a, b):
        s = a + b
        t = s - a
        return s, (a - (s - t)) + (b - t)

    def fast_two_sum(a, b):
        s = a + b
        return s, b - (s - a)

    rng = random.Random(23)
    cases = [(0.0, 7.0), (1.0, 1.0), (3.0, 3.0), (2.0**53 - 2, 2.0**53 - 1)]
    for _ in range(20000):
        y = rng.randrange(1, 1 << rng.choice((20, 33, 42, 48, 53)))
        cases.append((float(rng.randrange(0, y + 1)), float(y)))
    for x, y in cases:
        z1 = x / y
        p, e = _dekker_fused(-z1, y)
        z2 = ((p + x) + e) / y
        assert abs(z2) <= abs(z1), (x, y)
        assert two_sum(z1, z2) == fast_two_sum(z1, z2), (x, y)


def test_replay_step_numerator_fold():
    """The replay's dd_step divides r - m, one fma from the balanced residue x (r = x, or x + m
    when x < 0), where the reference divides m - r and negates the quotient while its sum is
    positive (Bellard.cl:1037-1048).  Every operation of nn_div and addcab is odd-symmetric under
    round to nearest, so the new sum is the same: equal doubles, a zero at most differing in
    sign.  Sums in (-1, 1) (cancelling ones too), |x| < m, integer moduli m < 2^53."""
    def two_sum(a, b):
        s = a + b
        t = s - a
        return s, (a - (s - t)) + (b - t)

    def nn_div(x, y):  # correctly rounded quotients, twoproduct as one fma, final Fast2Sum
        z1 = x / y
        p = -z1 * y
        e = fma(-z1, y, -p)
        z2 = ((p + x) + e) / y
        s = z1 + z2
        return s, z2 - (s - z1)

    def addcab(a, b):
        c = two_sum(a[0], b[0])
        return two_sum(c[0], (c[1] + a[1]) + b[1])

    def ref_step(s, x, m):
        r = x + m if x < 0 else x
        pos = s[0] > 0.0
        q = nn_div(m - r if pos else r, m)
        if pos:
            q = (-q[0], -q[1])
        return addcab(s, q)

    def folded_step(s, x, m):
        c = (1.0 if x < 0.0 else 0.0) - (1.0 if s[0] > 0.0 else 0.0)
        return addcab(s, nn_div(fma(c, m, x), m))

    rng = random.Random(29)
    seen = set()
    for i in range(30000):
        m = rng.randrange(1, 1 << rng.choice((20, 33, 42, 48, 53)))
        x = rng.randrange(-m + 1, m) if m > 1 else 0
        r = x + m if x < 0 else x
        kind = i % 4
        if kind == 0:  # a general sum in (-1, 1)
            hi = rng.uniform(-1.0, 1.0)
        elif kind == 1:  # one that the positive-branch quotient -(m - r) / m cancels
            hi = (m - r) / m * (1.0 + rng.choice((0.0, 2.0**-52, -(2.0**-52))))
        elif kind == 2:  # one that the non-positive-branch quotient r / m cancels
            hi = -(r / m)
        else:  # zero or tiny
            hi = rng.choice((0.0, -0.0, 2.0**-60, -(2.0**-70)))
        lo = rng.uniform(-0.5, 0.5) * math.ulp(hi) if hi != 0.0 else 0.0
        s = two_sum(hi, lo)
        a, b = ref_step(s, float(x), float(m)), folded_step(s, float(x), float(m))
        assert a == b, (s, x, m, a, b)
        seen.add((s[0] > 0.0, x < 0))
    assert len(seen) == 4


# ------------------------------------------------------------------ integer engine (plain kernel)

M64 = (1 << 64) - 1


def test_montgomery_one_carried_along_the_stride():
    """2^64 mod m carried from term to term (m += 64 den0) for m >= 2^37: one - qq s, plus m + s
    when negative, equals 2^64 mod m at every term of a tile and qq stays floor(2^64 / m)."""
    rng = random.Random(37)
    for _ in range(3000):
        den0 = rng.choice([4, 10])
        den1 = rng.choice([1, 3, 5, 7, 9]) if den0 == 10 else rng.choice([1, 3])
        k = rng.randrange(((1 << 37) - den1) // den0 + 1, 10**13)
        m = den0 * k + den1
        s = 64 * den0
        qq, one = (M64 // m), M64 - (M64 // m) * m + 1
        for _ in range(16):
            assert one == (1 << 64) % m and qq == (1 << 64) // m and qq < 1 << 27
            m += s
            t = qq * s
            assert t < m
            if one >= t:
                one -= t
            else:
                one, qq = one + m - t, qq - 1


def _redc_words(t0, t1, t2, t3, m, mp):
    """redc_words of clb_device.hpp, word by word with the kernel's 32/64-bit wraps."""
    m0, m1, mp0, mp1 = m & 0xFFFFFFFF, m >> 32, mp & 0xFFFFFFFF, mp >> 32
    qa = t0 * mp0
    q0 = qa & 0xFFFFFFFF
    q1 = ((qa >> 32) + t0 * mp1 + t1 * mp0) & 0xFFFFFFFF
    e = q0 * m0 + t0
    f = q1 * m0 + ((e >> 32) + t1)
    g = q0 * m1 + (f & 0xFFFFFFFF)
    hi = q1 * m1 + (f >> 32) + (g >> 32)
    assert max(qa, e, f, g, hi) <= M64
    assert e & 0xFFFFFFFF == 0 and g & 0xFFFFFFFF == 0  # the low words of q m + (t1:t0) vanish
    q = (q1 << 32) | q0
    assert hi == (q * m + ((t1 << 32) | t0)) >> 64
    r = ((t3 << 32) | t2) + hi
    assert r <= M64
    return r


def _isq55_shift(h, c, m, mp):
    gw = (h << c) & M64
    h0, h1, g0, g1 = h & 0xFFFFFFFF, h >> 32, gw & 0xFFFFFFFF, gw >> 32
    a = h0 * g0
    x = h0 * g1 + (a >> 32)
    x2 = h1 * g0 + (x & 0xFFFFFFFF)
    y = h1 * g1 + (x >> 32) + (x2 >> 32)
    assert max(a, x, x2, y) <= M64
    return _redc_words(a & 0xFFFFFFFF, x2 & 0xFFFFFFFF, y & 0xFFFFFFFF, y >> 32, m, mp)


def _isq55(h, m, mp):
    h0, h1 = h & 0xFFFFFFFF, h >> 32
    a = h0 * h0
    x = h0 * (h1 + h1) + (a >> 32)
    y = h1 * h1 + (x >> 32)
    assert max(a, x, y) <= M64
    return _redc_words(a & 0xFFFFFFFF, x & 0xFFFFFFFF, y & 0xFFFFFFFF, y >> 32, m, mp)


def test_montgomery_square_words_is_redc():
    """The integer engine's squares for m < 2^55 (chained 32x32+64 multiply-adds): with lazy
    residues h < 2m and window digits c <= 7, REDC(h^2 2^c) = h^2 2^c 2^-64 mod m (up to one m),
    the result stays below 2m, and no multiply-add leaves 64 bits."""
    rng = random.Random(55)
    for _ in range(20000):
        bits = rng.choice([3, 20, 33, 41, 42, 48, 54, 55])
        m = rng.randrange(1 << (bits - 1), 1 << bits) | 1
        if m >= 1 << 55:
            m -= 2
        mp = (-pow(m, -1, 1 << 64)) & M64
        h = rng.choice([rng.randrange(0, 2 * m), 2 * m - 1, 0, 1, m, m - 1])
        c = rng.randrange(0, 8)
        r = _isq55_shift(h, c, m, mp)
        assert r < 2 * m and r % m == (h * h << c) * pow(2, -64, m) % m, (m, h, c)
        r0 = _isq55(h, m, mp)
        assert r0 < 2 * m and r0 % m == h * h * pow(2, -64, m) % m, (m, h)

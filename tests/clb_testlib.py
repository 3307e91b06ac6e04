"""Shared test helpers: the CPU oracle (oracle/liboracle.so) and exact big-integer references.

The oracle is test infrastructure; the product library never loads it.
"""
from __future__ import annotations

import ctypes
import json
import os
import struct
import subprocess
from fractions import Fraction

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

M192 = (1 << 192) - 1
NUMER = (32, 1, 256, 64, 4, 4, 1)
NUMESIGN = (1, 1, 0, 1, 1, 1, 0)
DEN0 = (4, 4, 10, 10, 10, 10, 10)
DEN1 = (1, 3, 1, 3, 5, 7, 9)
SHIFT = (5, 0, 8, 6, 2, 2, 0)

U3 = ctypes.c_uint64 * 3
U4 = ctypes.c_uint64 * 4


class Oracle:
    """ctypes view of oracle/clb_oracle.h."""

    def __init__(self, path: str):
        lib = ctypes.CDLL(path)
        u64, i64, c_int = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int
        p64 = ctypes.POINTER(ctypes.c_uint64)
        sig = {
            "orc_normalize": ([u64, p64, p64, ctypes.POINTER(c_int)], c_int),
            "orc_formula": ([c_int, p64, ctypes.POINTER(i64), p64, p64], c_int),
            "orc_modinv32": ([ctypes.c_uint32], ctypes.c_uint32),
            "orc_modinv64": ([u64], u64),
            "orc_mr32": ([ctypes.c_uint32] * 4, ctypes.c_uint32),
            "orc_mr64": ([u64] * 4, u64),
            "orc_expmod32": ([u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32],
                             ctypes.c_uint32),
            "orc_expmod64": ([u64] * 4, u64),
            "orc_reciprocal_word": ([u64], u64),
            "orc_div2by1": ([u64, u64, u64, p64, u64], u64),
            "orc_divnby1": ([c_int, p64, u64, p64], None),
            "orc_mulmodn": ([u64, u64, u64, c_int, u64], u64),
            "orc_pow1024_mod": ([u64, u64], u64),
            "orc_term192": ([u64, u64, i64, u64, u64, p64], None),
            "orc_add192": ([p64, p64], None),
            "orc_tail192": ([u64, i64, u64, u64, p64], None),
            "orc_sum7_192": ([p64, p64, u64], None),
            "orc_range_sum": ([u64, u64, u64, c_int, c_int, p64], None),
            "orc_batch_calc": ([u64, u64, u64, c_int, c_int, p64], None),
            "orc_normal": ([u64, c_int, p64], c_int),
            "orc_hex48": ([p64, ctypes.c_char_p], None),
            "orc_hex_dd": ([ctypes.POINTER(ctypes.c_double), ctypes.c_char_p], None),
            "orc_work_sum_add": ([ctypes.POINTER(ctypes.c_int), ctypes.c_char_p], None),
            "orc_term256": ([u64, u64, i64, u64, u64, p64], None),
            "orc_tail256": ([u64, i64, u64, u64, p64], None),
            "orc_job256": ([u64, u64, u64, c_int, c_int, p64], None),
            "orc_normal256": ([u64, c_int, p64], c_int),
            "orc_ref_item192": ([u64, u64, u64, u64, u64, ctypes.c_int, U3], None),
            "orc_ref_tail192": ([u64, ctypes.c_int, U3], None),
            "orc_ref_batch192": ([u64, u64, u64, ctypes.c_int, ctypes.c_int, u64, ctypes.c_int,
                                  U3], ctypes.c_int),
            "orc_ref_normal192": ([u64, ctypes.c_int, u64, ctypes.c_int, U3], ctypes.c_int),
            "orc_ref_job106": ([u64, u64, u64, ctypes.c_int, u64, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "orc_ref_normal106": ([u64, u64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                                  ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        self.lib = lib
        self.threads = max(1, min(16, os.cpu_count() or 1))

    # -- wrappers returning Python ints (192-bit values as integers)
    def normalize(self, n):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        rc = self.lib.orc_normalize(n, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return None if rc else (a.value, b.value, bool(c.value))

    def term192(self, d, k, numesign, den0, den1):
        q = U3()
        self.lib.orc_term192(d, k, numesign, den0, den1, q)
        return w2i(q)

    def tail192(self, d, numesign, den0, den1):
        q = U3()
        self.lib.orc_tail192(d, numesign, den0, den1, q)
        return w2i(q)

    def range_sum(self, d, k0, k1, j, threads=None):
        q = U3()
        self.lib.orc_range_sum(d, k0, k1, j, threads or self.threads, q)
        return w2i(q)

    def batch_calc(self, d, k0, k1, j, threads=None):
        q = U3()
        self.lib.orc_batch_calc(d, k0, k1, j, threads or self.threads, q)
        return w2i(q)

    def normal(self, n, threads=None):
        q = U3()
        rc = self.lib.orc_normal(n, threads or self.threads, q)
        assert rc == 0
        return w2i(q)

    # -- the 256-bit error oracle (values as integers: floor(value * 2^256))
    def term256(self, d, k, numesign, den0, den1):
        q = U4()
        self.lib.orc_term256(d, k, numesign, den0, den1, q)
        return w4i(q)

    def tail256(self, d, numesign, den0, den1):
        q = U4()
        self.lib.orc_tail256(d, numesign, den0, den1, q)
        return w4i(q)

    def job256(self, d, k0, k1, j, threads=None):
        """c_j * (formula j over [k0, min(k1, d)), plus the tail iff it reaches d) mod 1."""
        q = U4()
        self.lib.orc_job256(d, k0, k1, j, threads or self.threads, q)
        return w4i(q)

    def normal256(self, n, threads=None):
        q = U4()
        assert self.lib.orc_normal256(n, threads or self.threads, q) == 0
        return w4i(q)

    def sum7(self, ans, part, numer):
        a, p = U3(*i2w(ans)), U3(*i2w(part))
        self.lib.orc_sum7_192(a, p, numer)
        return w2i(a), w2i(p)

    def hex48(self, v):
        buf = ctypes.create_string_buffer(49)
        self.lib.orc_hex48(U3(*i2w(v)), buf)
        return buf.value.decode()

    def hex_dd(self, hi, lo):
        buf = ctypes.create_string_buffer(49)
        self.lib.orc_hex_dd((ctypes.c_double * 2)(hi, lo), buf)
        return buf.value.decode()

    def pow1024_mod(self, n, m):
        return self.lib.orc_pow1024_mod(n, m)

    # -- replay of the reference's precision-1/2 kernels (double-double rounding included)
    def ref_item192(self, d, offset, k_max, idx, gsize, j):
        q = U3()
        self.lib.orc_ref_item192(d, offset, k_max, idx, gsize, j, q)
        return w2i(q)

    def ref_tail192(self, d, j):
        q = U3()
        self.lib.orc_ref_tail192(d, j, q)
        return w2i(q)

    def ref_batch192(self, d, k0, k1, j, precision, gsize=1 << 24, threads=None):
        q = U3()
        rc = self.lib.orc_ref_batch192(d, k0, k1, j, precision, gsize, threads or self.threads, q)
        assert rc == 0
        return w2i(q)

    def ref_job106(self, d, k0, k1, j, gsize=1 << 24, threads=None):
        """(hi, lo) of bigSum[0] after BatchCalc and Sum7 in the 106-bit mode."""
        out = (ctypes.c_double * 2)()
        rc = self.lib.orc_ref_job106(d, k0, k1, j, gsize, threads or self.threads, out)
        assert rc == 0
        return out[0], out[1]

    def ref_normal106(self, n, gsize=1 << 24, threads=None):
        out = (ctypes.c_double * 2)()
        rc = self.lib.orc_ref_normal106(n, gsize, threads or self.threads, out)
        assert rc == 0
        return out[0], out[1]

    def ref_normal192(self, n, precision, gsize=1 << 24, threads=None):
        q = U3()
        rc = self.lib.orc_ref_normal192(n, precision, gsize, threads or self.threads, q)
        assert rc == 0
        return w2i(q)


def w2i(w) -> int:
    return (int(w[0]) << 128) | (int(w[1]) << 64) | int(w[2])


def w4i(w) -> int:
    return (int(w[0]) << 192) | (int(w[1]) << 128) | (int(w[2]) << 64) | int(w[3])


M256 = (1 << 256) - 1


def frac_error_256(result192: int, exact256: int) -> int:
    """Signed error of a 192-bit fraction against a 256-bit one, in units of 2^-256 (the
    difference taken on the unit circle, in (-2^255, 2^255])."""
    e = ((result192 << 64) - exact256) & M256
    return e - (1 << 256) if e > (1 << 255) else e


def error_log2(units256: int) -> float:
    """log2 of |error| as a fraction (units of 2^-256); -inf for 0."""
    import math
    a = abs(units256)
    return -math.inf if a == 0 else math.log2(a) - 256


def i2w(v: int):
    v &= M192
    return (v >> 128, (v >> 64) & (2**64 - 1), v & (2**64 - 1))


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load_oracle() -> Oracle:
    src = os.path.join(ORACLE_DIR, "clb_oracle.c")
    if not os.path.exists(ORACLE_LIB) or os.path.getmtime(ORACLE_LIB) < os.path.getmtime(src):
        build_oracle()
    return Oracle(ORACLE_LIB)


def load_pi_fixtures():
    with open(os.path.join(GOLDEN, "pi_hex.json")) as f:
        return json.load(f)


# ----------------------------------------------------------------- exact big-integer references

def py_normalize(n: int):
    d = ((n - 1) * 2 - 3) // 5
    if d % 2 == 0:
        d -= 1
    return (d * 5 + 3) // 2 + 1, d


def py_term(d: int, k: int, numesign: int, den0: int, den1: int) -> int:
    """floor(2^192 frac(+-1024^(d-k)/m)) for a head term k < d."""
    m = den0 * k + den1
    r = pow(2, 10 * (d - k), m)
    if (numesign + k) % 2 == 1:
        r = m - r
    return ((r << 192) // m) & M192


def py_tail_term(d: int, i: int, numesign: int, den0: int, den1: int) -> int:
    k = d + i
    m = den0 * k + den1
    q = (1 << (192 - 10 * i)) // m
    if (numesign + k) % 2 == 1:
        q = M192 - q  # floor((1 - x) 2^192) for the non-integer x 2^192
    return q


def py_term256(d: int, k: int, numesign: int, den0: int, den1: int) -> int:
    m = den0 * k + den1
    r = pow(2, 10 * (d - k), m)
    if (numesign + k) % 2 == 1:
        r = m - r
    return ((r << 256) // m) & M256


def py_tail256(d: int, j: int) -> int:
    s = 0
    for i in range(26):
        k = d + i
        m = DEN0[j] * k + DEN1[j]
        q = (1 << (256 - 10 * i)) // m
        if (NUMESIGN[j] + k) % 2 == 1:
            q = M256 - q
        s += q
    return s & M256


def py_range_sum(d: int, k0: int, k1: int, j: int) -> int:
    s = 0
    for k in range(k0, k1):
        s += py_term(d, k, NUMESIGN[j], DEN0[j], DEN1[j])
    return s & M192


def py_tail(d: int, j: int) -> int:
    return sum(py_tail_term(d, i, NUMESIGN[j], DEN0[j], DEN1[j]) for i in range(20)) & M192


def py_batch_calc(d: int, k_start: int, k_end: int, j: int) -> int:
    hi = min(k_end, d)
    offset = k_start
    s = 0
    if k_start < hi:
        s = py_range_sum(d, k_start, hi, j)
        offset = hi
    if offset == d:
        s += py_tail(d, j)
    return s & M192


def py_normal(n: int) -> int:
    _, d = py_normalize(n)
    ans = 0
    for j in range(7):
        ans += (py_batch_calc(d, 0, d, j) << SHIFT[j])
    return ans & M192


def bound_ulps(d: int) -> int:
    """Truncation loss of the exact-integer semantics: each term's floor loses < 1 unit, scaled
    by c_j in the combine; the tail beyond 20 terms is far below one unit."""
    return sum(NUMER) * (d + 20) + 2


def pi_series_check(terms: int = 12) -> Fraction:
    """Bellard's series with the formula table, as an exact rational (for pinning the table)."""
    total = Fraction(0)
    for n in range(terms):
        s = Fraction(0)
        for j in range(7):
            sign = -1 if NUMESIGN[j] else 1
            s += sign * Fraction(NUMER[j], DEN0[j] * n + DEN1[j])
        total += Fraction((-1) ** n, 1024 ** n) * s
    return total / 64


# ----------------------------------------- the reference's own kernels (oracle/_ref, GPU only)

REF_DIR = os.path.join(ORACLE_DIR, "_ref")
REF_LIB = os.path.join(REF_DIR, "libref_runner.so")


def loopmax32(j: int) -> int:
    """First k whose modulus leaves 32 bits (KernelFuncs.as:31)."""
    return ((1 << 32) - DEN1[j] - 1) // DEN0[j] + 1


# Term ranges for the double-double replay of the reference's precision-1/2 kernels, as
# (d, k_start, k_end, j, precision, work-items per launch).  The reference launches 2^24
# work-items; smaller launches (its kernels stride by the global size) give every work-item its
# full 16 or 32 terms on ranges the CPU restatement finishes in seconds.
REF_DD_GOLDEN = os.path.join(GOLDEN, "ref_dd192.json")
REF_DD106_GOLDEN = os.path.join(GOLDEN, "ref_dd106.json")
REF_OD_GOLDEN = os.path.join(GOLDEN, "ref_dd_overdouble.json")


def ref_dd_cases():
    d6, d7 = py_normalize(10**6)[1], py_normalize(10**7)[1]          # 399_999, 3_999_999
    d10, d12 = py_normalize(25 * 10**9)[1], py_normalize(10**12)[1]  # 9_999_999_999, 4e11 - 1
    cases = []
    for j in range(7):
        for p in (1, 2):
            cases.append((d6, 0, d6, j, p, 1 << 24))  # one term per work-item, and the tail
            cases.append((d6, 0, d6, j, p, 4096))     # 16 terms per work-item
    for j in (0, 3, 6):
        cases.append((d7, 0, d7, j, 1, 65536))
        cases.append((d7, 12345, d7 // 2 + 7, j, 2, 2048))  # odd start, no tail
    for j in range(7):
        lo = loopmax32(j)
        for p in (1, 2):
            cases.append((d10, lo - 3001, lo + 9000, j, p, 256))  # both head kernels
    for j in (1, 4):
        cases.append((d10, d10 - 5000, d10, j, 1, 256))           # 64-bit kernel and the tail
        cases.append((d10, d10 - 5000, d10, j, 2, 256))           # Refresh and the tail
        cases.append((d12, d12 - 4000, d12 + 7, j, 1, 512))       # k_end past d
        cases.append((d12, d12 // 2 | 1, d12 // 2 + 8192, j, 1, 256))
        cases.append((d12, 0, 6000, j, 1, 256))                   # the smallest moduli
    cases.append((d6, d6, d6, 2, 1, 1 << 24))                     # the tail alone
    return cases


def ref_dd106_cases():
    """106-bit mode jobs (d, k_start, k_end, j, work-items per launch), small enough for the CPU
    restatement's literal 65536-group trees."""
    d6, d7 = py_normalize(10**6)[1], py_normalize(10**7)[1]
    d10, d12 = py_normalize(25 * 10**9)[1], py_normalize(10**12)[1]
    cases = []
    for j in range(7):
        cases.append((d6, 0, d6, j, 1 << 24))   # one launch, the tail
        cases.append((d6, 0, d6, j, 4096))      # 16 terms per work-item, 7 launches
    for j in (0, 5):
        cases.append((d7, 777, d7 // 3, j, 2048))
        lo = loopmax32(j)
        cases.append((d10, lo - 3001, lo + 9000, j, 256))     # Sglobal32mtg, then Sglobal64mtg
        cases.append((d10, d10 - 5000, d10 + 3, j, 512))      # Sglobal64mtg and the tail
        cases.append((d12, d12 // 2 | 1, d12 // 2 + 8192, j, 256))
    return cases


def ref_od_cases():
    """Jobs (d, k_start, k_end, j, precision, work-items per launch) at loop counts with
    10 d + 9 >= 2^53, where the reference host builds its kernels with -D OVERDOUBLE
    (Bellard_192.hsp:472-474), in all three precisions: moduli below 2^53 (the smallest such d),
    across 2^53 inside one launch, above it (den0 4 and 10), and above 2^55 (the wide engine)."""
    d_a = 900719925474099           # the smallest: 10 d + 9 >= 2^53, every modulus below 2^53
    d_b = 1000000000000001
    d_c = 3000000000000001
    d_d = 5000000000000001          # moduli up to 5e16 > 2^55
    k53 = (1 << 53) // 10           # den0 = 10: m crosses 2^53 at k = k53
    cases = []
    for p in (0, 1, 2):
        cases += [
            (d_a, d_a - 4000, d_a, 2, p, 256),              # the tail, moduli below 2^53
            (d_b, 0, 3000, 0, p, 256),                      # Sglobal32mtg alone
            (d_b, loopmax32(3) - 1500, loopmax32(3) + 2500, 3, p, 256),  # both head kernels
            (d_b, k53 - 1000, k53 + 2000, 4, p, 256),       # across 2^53
            (d_b, d_b - 4000, d_b, 6, p, 256),              # above 2^53, and the tail
            (d_c, d_c - 3000, d_c, 0, p, 512),              # den0 4 above 2^53
            (d_d, d_d - 2500, d_d + 5, 5, p, 256),          # above 2^55, k_end past d
        ]
    return cases


class RefKernels:
    """ctypes view of oracle/ref_runner.c: the reference's OpenCL kernels (compiled from its
    sources into oracle/_ref/) driven through the OpenCL runtime with the reference's launch
    sequence.  precision 0 returns the double-double (hi, lo); 1 and 2 return 192-bit words in
    bigSum order (most significant first)."""

    def __init__(self, path: str = REF_LIB, co_dir: str = REF_DIR):
        lib = ctypes.CDLL(path)
        u64, c_int = ctypes.c_uint64, ctypes.c_int
        p64 = ctypes.POINTER(ctypes.c_uint64)
        lib.ref_init.argtypes, lib.ref_init.restype = [ctypes.c_char_p], c_int
        lib.ref_normal.argtypes, lib.ref_normal.restype = [u64, c_int, p64], c_int
        lib.ref_job.argtypes = [u64, u64, u64, c_int, c_int, p64]
        lib.ref_job.restype = c_int
        lib.ref_last_error.restype = ctypes.c_char_p
        lib.ref_last_ms.restype = ctypes.c_double
        lib.ref_shutdown.restype = None
        lib.ref_set_gws.argtypes, lib.ref_set_gws.restype = [u64], c_int
        lib.ref_op.argtypes = [c_int, c_int, u64, u64, u64, c_int, p64]
        lib.ref_op.restype = c_int
        self.lib = lib
        self._check(lib.ref_init(co_dir.encode()))

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise RuntimeError("reference kernels: " + self.lib.ref_last_error().decode())

    @staticmethod
    def _value(w, precision: int):
        if precision == 0:
            hi, lo = (struct.unpack("<d", struct.pack("<Q", x))[0] for x in w[:2])
            return hi, lo
        return tuple(w)

    def normal(self, d: int, precision: int):
        w = U3()
        self._check(self.lib.ref_normal(d, precision, w))
        return self._value(w, precision)

    def job(self, d: int, k0: int, k1: int, j: int, precision: int):
        w = U3()
        self._check(self.lib.ref_job(d, k0, k1, j, precision, w))
        return self._value(w, precision)

    def op(self, op: int, precision: int, a: int = 0, b: int = 0, c: int = 0, j: int = 0):
        """One reference kernel on bigSumMem, then bigSum[0]: op 0 Set0mem, 1 Sglobal32mtg*(offset
        a, d b, k_max c), 2 Sglobal64mtg*(a, b, c), 3 Sglobal_after*(d b), 4 Ssum's two passes."""
        w = U3()
        self._check(self.lib.ref_op(op, precision, a, b, c, j, w))
        return self._value(w, precision)

    def set_gws(self, gws: int = 1 << 24) -> None:
        """Work-items per term-kernel launch (the reference launches 2^24)."""
        self._check(self.lib.ref_set_gws(gws))

    def last_ms(self) -> float:
        return self.lib.ref_last_ms()

    def close(self) -> None:
        self.lib.ref_shutdown()


def dd_error_ulps(hi: float, lo: float, exact: int) -> int:
    """|frac(hi + lo) - exact / 2^192| in units of 2^-192 (distance on the unit circle)."""
    v = (Fraction(hi) + Fraction(lo)) * (1 << 192)
    diff = (v - exact) % (1 << 192)
    diff = min(diff, (1 << 192) - diff)
    return int(diff) + (1 if diff.denominator != 1 else 0)


def u192_error_ulps(w, exact: int) -> int:
    diff = (w2i(w) - exact) % (1 << 192)
    return min(diff, (1 << 192) - diff)


# ----------------------------------------------------------------- seeded synthetic workloads

def splitmix64(seed: int):
    """SplitMix64 (Steele, Lea, Flood 2014): the explicit generator of every seeded workload."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def batched_positions(count: int = 4096, seed: int = 0):
    """Digit positions log-uniform in [10^6, 10^11], multiples of 5: u = (x >> 11) 2^-53,
    N = 5 round(10^(6 + 5u) / 5)."""
    rng = splitmix64(seed)
    out = []
    for _ in range(count):
        u = (next(rng) >> 11) * 2.0**-53
        out.append(5 * round(10 ** (6 + 5 * u) / 5))
    return out


def uneven_jobs(d: int, count: int = 64, seed: int = 1):
    """Term-range cut points of `count` jobs whose lengths are drawn uniformly from
    [10^8, 10^11] and rescaled so that the jobs tile [0, d) exactly: c_i = round(d sum_{<=i} L /
    sum L).  Returns the count + 1 cut points (0 first, d last)."""
    rng = splitmix64(seed)
    lens = [10**8 + next(rng) % (10**11 - 10**8 + 1) for _ in range(count)]
    total, acc, cuts = sum(lens), 0, [0]
    for ln in lens:
        acc += ln
        cuts.append((d * acc + total // 2) // total)
    return cuts

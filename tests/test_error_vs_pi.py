"""GPU: the fraction error of every mode measured against pi itself.

At each fixture position the exact 256-bit value floor(frac(16^(N-1) pi) 2^256) comes from
mpmath (tests/golden/pi_hex.json, "digits256").  The exact mode (both term kernels) must lie below
it by at most its per-term floor losses (one unit of 2^-192 per term times c_j); the replay of the
reference's kernels (precision 2, 1: 192-bit, precision 0: double-double) within their
double-double bounds; and the measured errors are printed.
"""
import math
import struct
import sys
from fractions import Fraction

import pytest

import clbellard_amd as clb
from clb_testlib import NUMER, bound_ulps, frac_error_256, py_normalize, w2i

pytestmark = pytest.mark.gpu

POSITIONS = [1000, 65540, 100000, 250000, 500000, 777775, 1000000]


def lg(units256):
    return "0" if units256 == 0 else f"2^{math.log2(abs(units256)) - 256:.1f}"


@pytest.mark.parametrize("n", POSITIONS)
def test_measured_error_against_pi(pi_fixtures, n):
    _, d = py_normalize(n)
    pi256 = int(pi_fixtures["digits256"][str(n)], 16)
    errs = {}
    for kernel in (clb.KERNEL_SPLIT, clb.KERNEL_PLAIN):
        with clb.Context(0, clb.PREC_U192, kernel=kernel) as c:
            errs[f"exact{kernel}"] = frac_error_256(w2i(c.normal(n)[0]), pi256)
    assert errs["exact1"] == errs["exact2"]
    assert -(bound_ulps(d) << 64) <= errs["exact2"] <= 0
    dd192 = sum(NUMER) * (2 * (d + 20) + (1 << 16)) << 64
    for p in (clb.PREC_U192_BETA, clb.PREC_U192):
        with clb.Context(0, p) as c:
            c.set_replay(clb.REF_GLOBAL_WORK_SIZE)
            errs[f"replay{p}"] = frac_error_256(w2i(c.normal(n)[0]), pi256)
        assert abs(errs[f"replay{p}"]) <= dd192 + (bound_ulps(d) << 64), (n, p)
    with clb.Context(0, clb.PREC_DD106) as c:
        c.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        frac = c.normal(n)[0]
        hi, lo = (struct.unpack("<d", struct.pack("<Q", w))[0] for w in frac[:2])
        v = int(((Fraction(hi) + Fraction(lo)) % 1) * (1 << 256))
        e = (v - pi256) % (1 << 256)
        errs["replay0"] = e - (1 << 256) if e > (1 << 255) else e
    assert abs(errs["replay0"]) <= (7 * (d + 20)) << (256 - 104), n
    print(f"N={n}: error vs pi: exact {lg(errs['exact2'])}, replay precision 2 "
          f"{lg(errs['replay2'])}, 1 {lg(errs['replay1'])}, 0 {lg(errs['replay0'])}",
          file=sys.stderr, flush=True)

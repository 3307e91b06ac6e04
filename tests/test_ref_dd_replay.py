"""The double-double replay of the reference's precision-1/2 kernels against those kernels.

The reference's 192-bit kernels round: each work-item keeps two double-double sums over its
strided terms (Bellard.cl:771-823, :992-1084) and converts them once (dd2_to_ulong3,
:506-578); the tail kernel is double-double throughout (:1153-1253).  orc_ref_batch192 restates
that arithmetic operation by operation.  Here it is held, bit for bit, to the reference's own
kernels (oracle/_ref, GPU) on every case of ref_dd_cases().
"""
import os
import sys

import pytest

from clb_testlib import (M192, REF_LIB, SHIFT, RefKernels, load_oracle, ref_dd_cases, w2i)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REF_LIB),
                                 reason="oracle/_ref not built (needs the reference sources)")]

CASES = ref_dd_cases()


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def ref():
    note("reference kernels: loading oracle/_ref")
    r = RefKernels()
    yield r
    r.set_gws()
    r.close()


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d{}_k{}-{}_j{}_p{}_g{}".format(*c))
def test_oracle_replay_equals_reference_kernels(ref, orc, case):
    d, k0, k1, j, p, gws = case
    note(f"reference kernels d={d} k=[{k0},{k1}) j={j} precision={p} gws={gws}")
    ref.set_gws(gws)
    got = w2i(ref.job(d, k0, k1, j, p))
    want = (orc.ref_batch192(d, k0, k1, j, p, gws) << SHIFT[j]) & M192
    assert got == want, f"{got:048x} != {want:048x}"

"""GPU tests against the reference's own kernels.

oracle/Makefile compiles Bellard.cl and Bellard_div2by1.cl from the reference sources into gfx950
code objects under oracle/_ref/; oracle/ref_runner.c replays the reference host's launch sequence
on them through the OpenCL runtime (KernelFuncs.as:25-74, Bellard_192.hsp:264-333).  These tests
hold our HIP path to them:

* precision 2's 64-bit kernel Sglobal64mtg_192_Refresh (Bellard_div2by1.cl:313-362) evaluates
  every term exactly, as ours do, so on term ranges past loopmax32 (where it is the only head
  kernel) and short of d (the tail kernel is double-double) the sums must be bit-identical;
* the double-double kernels (every precision below loopmax32, precision 1 and 0 above it, the
  tails) carry rounding error; their distance from our exact sums is measured, bounded and
  printed, and the first 24 hex digits of a position (16 in the 106-bit mode) must agree;
* the reference kernels' rate on the same GPU is printed beside ours.
"""
import os
import sys
import time

import pytest

import clbellard_amd as clb
from clb_testlib import (NUMER, REF_LIB, RefKernels, dd_error_ulps, i2w, loopmax32,
                         py_normalize, u192_error_ulps, w2i)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REF_LIB),
                                 reason="oracle/_ref not built (needs the reference sources)")]


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def log2(x: int) -> float:
    return float("-inf") if x == 0 else x.bit_length() - 1 + (x / (1 << (x.bit_length() - 1)) - 1)


@pytest.fixture(scope="module")
def ref():
    note("reference kernels: loading oracle/_ref")
    r = RefKernels()
    yield r
    r.close()


@pytest.fixture(scope="module")
def ctx():
    c = clb.Context(0, clb.PREC_U192)
    yield c
    c.close()


D10 = py_normalize(25_000_000_000)[1]          # 9_999_999_999: both head kernels run
D12 = py_normalize(1_000_000_000_000)[1]       # 399_999_999_999: the headline position


def refresh_ranges(d, j):
    """Term ranges that only Sglobal64mtg_192_Refresh covers: from loopmax32, an odd start in the
    middle, and one ending just short of d (small exponents)."""
    lo = loopmax32(j)
    mid = d // 2 | 1
    return [(lo, lo + (1 << 25) + 12345), (mid, mid + 3 * (1 << 20) + 7),
            (d - (1 << 22) - 3, d - 1)]


@pytest.mark.parametrize("j", range(7))
def test_refresh_kernel_bit_exact(ref, ctx, j):
    for k0, k1 in refresh_ranges(D10, j):
        note(f"Refresh j={j} d={D10} k=[{k0},{k1})")
        want = ctx.run_jobs([(D10, k0, k1, j)])[0]
        got = ref.job(D10, k0, k1, j, 2)
        assert got == want, (j, k0, k1, f"{w2i(got):048x}", f"{w2i(want):048x}")


def test_refresh_kernel_bit_exact_headline_position(ref, ctx):
    """d = 399_999_999_999 (N = 10^12): one full reference launch (2^29 terms) per formula."""
    for j in range(7):
        k0 = D12 // 3 + 1000 * j
        note(f"Refresh j={j} d={D12} k=[{k0},{k0 + (1 << 29)})")
        want = ctx.run_jobs([(D12, k0, k0 + (1 << 29), j)])[0]
        got = ref.job(D12, k0, k0 + (1 << 29), j, 2)
        assert got == want, j


# Error bounds of the double-double kernels against the exact sums, in units of 2^-192 before the
# c_j scaling.  Each term of Sglobal32mtg_192 / Sglobal64mtg_192 is two double-double quotients
# (Bellard.cl:771-822, :992-1090) joined at bit 96 by dd2_to_ulong3 (Bellard.cl:506-580); the
# join truncates, so the error grows linearly with the term count (measured on MI355X: about
# 2^-6.6 units per term for the 32-bit kernel, 2^-2 for the 64-bit one).  The bound allows 2 units
# per term, plus 2^16 for the tail kernel Sglobal_after_192.
def dd192_bound(terms: int, numer: int = 1) -> int:
    return numer * (2 * terms + (1 << 16))


@pytest.mark.parametrize("precision", [1, 2])
def test_double_double_kernels_error(ref, ctx, precision):
    cases = [(D10, 0, 1 << 24, j) for j in range(7)]                        # 32-bit kernel
    cases += [(D10, D10 - (1 << 22), D10 + 5, j) for j in range(7)]        # reaches the tail
    if precision == 1:
        cases += [(D10, k0, k1, j) for j in range(7) for k0, k1 in refresh_ranges(D10, j)[:2]]
    worst = 0
    for d, k0, k1, j in cases:
        want = ctx.run_jobs([(d, k0, k1, j)])[0]
        got = ref.job(d, k0, k1, j, precision)
        err = u192_error_ulps(got, w2i(want))
        terms = min(k1, d) - k0
        note(f"precision {precision} j={j} k=[{k0},{k1}) error 2^{log2(err):.1f} ulp192")
        worst = max(worst, err)
        assert err <= dd192_bound(terms, NUMER[j]), (precision, j, k0, k1, err)
    note(f"precision {precision}: worst error 2^{log2(worst):.1f} x 2^-192")


POSITIONS = [1_000_000, 10_000_000, 100_000_000, 1_000_000_000, 2_000_000_000, 10_000_000_000]


@pytest.mark.parametrize("n", POSITIONS)
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_positions_match_reference_kernels(ref, ctx, n, precision):
    """Normal mode (Allloop, Bellard_192.hsp:264-296) on the reference kernels against ours: the
    first 24 hex digits agree, the fractions differ by the double-double rounding only."""
    nn, d = py_normalize(n)
    note(f"position N={nn} precision {precision}")
    frac, hexs, _ = ctx.normal(n)
    exact = w2i(frac)
    t0 = time.perf_counter()
    v = ref.normal(d, precision)
    ref_s = time.perf_counter() - t0
    if precision == 0:
        err = dd_error_ulps(v[0], v[1], exact)
        # 106-bit accumulation of 7 (d + 10) quotients, each addition within 2^-104 (worst case
        # linear in the term count; the measured error is far smaller)
        bound = 7 * (d + 20) << (192 - 104)
        ref_hex = clb.hex_dd(*v)
        ours_hex = clb.hex_dd(*clb.u192_to_dd(frac))
        ndig = 16
    else:
        err = u192_error_ulps(v, exact)
        bound = dd192_bound(d + 20, sum(NUMER))
        ref_hex = clb.hex48(v)
        ours_hex = hexs
        ndig = 24
    note(f"  ours {hexs}\n  ref  {ref_hex}  error 2^{log2(err):.1f} x 2^-192, "
         f"reference {ref_s:.2f} s")
    assert ref_hex[:ndig] == ours_hex[:ndig]
    assert err <= bound, (n, precision, err)


def test_reference_kernel_rate(ref, ctx):
    """The reference's head kernels against ours on the same GPU at N = 10^12: three windows of
    2^29 terms (one reference launch each) of formula 2 (m = 10k + 1)."""
    wins = [(D12, D12 // 4 * i + 1, D12 // 4 * i + 1 + (1 << 29), 2) for i in (1, 2, 3)]
    terms = 3 << 29
    ctx.run_jobs(wins)  # warm-up
    want = ctx.run_jobs(wins)
    ours_ms = ctx.last_timing()[0]
    rates = {"ours": terms / ours_ms * 1e3}
    for precision in (2, 1):
        got = []
        ms = 0.0
        for w in wins:
            got.append(ref.job(*w, precision))
            ms += ref.last_ms()
        rates[f"reference precision {precision}"] = terms / ms * 1e3
        if precision == 2:
            assert got == want
    for k, v in rates.items():
        note(f"rate {k}: {v:.3e} term-formulas/s")
    assert rates["ours"] > rates["reference precision 2"]
    assert rates["ours"] > rates["reference precision 1"]


@pytest.mark.skipif(not os.environ.get("CLB_REF_FULL_N"),
                    reason="a full reference run takes minutes; set CLB_REF_FULL_N=<position>")
@pytest.mark.parametrize("precision", [2, 1])
def test_full_position_reference_kernels(ref, ctx, precision):
    """A whole position on the reference kernels, formula by formula (Allloop,
    Bellard_192.hsp:264-296; dev_ans accumulation done here as the mod 2^192 sum of the Sum7
    outputs), against ours: 24 hex digits equal, error within the double-double bound, and both
    wall times printed."""
    n = int(os.environ["CLB_REF_FULL_N"])
    nn, d = py_normalize(n)
    t0 = time.perf_counter()
    frac, hexs, _ = ctx.normal(n)
    ours_s = time.perf_counter() - t0
    note(f"ours N={nn}: {hexs} in {ours_s:.2f} s")
    total, ref_s = 0, 0.0
    for j in range(7):
        t0 = time.perf_counter()
        part = ref.job(d, 0, d, j, precision)
        dt = time.perf_counter() - t0
        ref_s += dt
        total = (total + w2i(part)) % (1 << 192)
        note(f"reference precision {precision} formula {j}: {dt:.2f} s")
    ref_hex = f"{total:048x}"
    err = u192_error_ulps(i2w(total), w2i(frac))
    note(f"reference N={nn} precision {precision}: {ref_hex} in {ref_s:.2f} s, "
         f"error 2^{log2(err):.1f} x 2^-192; {7 * (d + 20) / ref_s:.3e} vs "
         f"{7 * (d + 20) / ours_s:.3e} term-formulas/s")
    assert ref_hex[:24] == hexs[:24]
    assert err <= dd192_bound(d + 20, sum(NUMER))


def test_work_files_match_reference_kernels(ref, tmp_path):
    """Work mode (Batchloop + fsave, Bellard_192.hsp:298-333) in precision 2 against the
    reference kernels' job results (Set0mem, BatchCalc, AddAns, then bigSum as fsave writes it):
    every job file whose range only Sglobal64mtg_192_Refresh covers is identical; the others are
    within the double-double bound; the files sum (workmodeSum) to the normal-mode digits."""
    n = 25_000_000_000
    _, d = py_normalize(n)
    batch = 1 << 30
    nb = d // batch + 1
    wl = tmp_path / "worklist.txt"
    lines = [f"digits={n}", "precision=2"] + [f"{j}_{b}" for j in range(7) for b in range(nb)]
    wl.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    work = str(tmp_path / "work")
    with clb.Context(0, clb.PREC_U192_BETA) as c:
        assert c.run_worklist(str(wl), work, batch_size=batch) == 7 * nb
        digits = c.normal(n)[1]
    exact = 0
    for j in range(7):
        for b in range(nb):
            k0, k1 = b * batch, (b + 1) * batch
            line0 = open(os.path.join(work, str(j), f"{j}_{b}.txt"), "rb").read().split(b"\r\n")[0]
            ours = int(line0, 16)
            theirs = w2i(ref.job(d, k0, k1, j, 2))
            if k0 >= loopmax32(j) and k1 < d:
                assert ours == theirs, (j, b)
                exact += 1
            else:
                err = min((ours - theirs) % (1 << 192), (theirs - ours) % (1 << 192))
                assert err <= dd192_bound(min(k1, d + 20) - k0, NUMER[j]), (j, b, err)
    note(f"{exact} of {7 * nb} work files bit-identical to the reference kernels' jobs")
    assert exact >= 7 * (nb - 3)
    assert clb.work_sum(work) == (digits, [nb] * 7)

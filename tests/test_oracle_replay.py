"""CPU: the double-double replay of the reference's kernels (orc_ref_batch192 for precision 1/2,
orc_ref_job106 / orc_ref_normal106 for the 106-bit mode) against the reference kernels' own
results in tests/golden/ref_dd192.json and ref_dd106.json (written by
tests/golden/make_ref_dd_fixtures.py on an MI355X with oracle/_ref) and, for loop counts with
10 d + 9 >= 2^53 (the -D OVERDOUBLE build), ref_dd_overdouble.json (make_ref_od_fixtures.py), bit
for bit, and the replay's distance from the exact sums within the double-double error the
reference carries."""
import json

import pytest

from clb_testlib import (M192, REF_DD106_GOLDEN, REF_DD_GOLDEN, REF_OD_GOLDEN, SHIFT, load_oracle,
                         py_batch_calc, py_normalize, ref_dd106_cases, ref_dd_cases, ref_od_cases)


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def golden():
    with open(REF_DD_GOLDEN) as f:
        return json.load(f)["cases"]


def test_fixture_covers_the_cases():
    have = {(c["d"], c["k0"], c["k1"], c["j"], c["precision"], c["gws"]) for c in golden()}
    assert have == set(ref_dd_cases())


def case_id(c):
    return "d{d}_k{k0}-{k1}_j{j}_p{precision}_g{gws}".format(**c)


@pytest.mark.parametrize("case", golden(), ids=case_id)
def test_replay_equals_reference_kernels(orc, case):
    d, k0, k1, j, p, gws = (case[k] for k in ("d", "k0", "k1", "j", "precision", "gws"))
    want = int(case["part"], 16)
    got = (orc.ref_batch192(d, k0, k1, j, p, gws) << SHIFT[j]) & M192
    assert got == want, f"{got:048x} != {want:048x}"


@pytest.mark.parametrize("gws", [1 << 24, 4096, 256])
def test_replay_close_to_exact(orc, gws):
    """Each work-item's conversion truncates at bit 192 and its sums round near bit 202, so the
    replay lies within 2 units of 2^-192 per term of the exact sum (measured: about 0.5 per term
    when work-items hold 16 terms, 0 for most single-term work-items)."""
    d = py_normalize(10**5)[1]
    for j in range(7):
        exact = py_batch_calc(d, 0, d, j)
        diff = (orc.ref_batch192(d, 0, d, j, 1, gws) - exact) & M192
        diff = diff - (1 << 192) if diff >> 191 else diff
        assert abs(diff) <= 2 * (d + 20), (j, diff)


def golden106():
    with open(REF_DD106_GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("case", golden106()["cases"],
                         ids=lambda c: "d{d}_k{k0}-{k1}_j{j}_g{gws}".format(**c))
def test_106bit_replay_equals_reference_kernels(orc, case):
    want = (float.fromhex(case["hi"]), float.fromhex(case["lo"]))
    assert orc.ref_job106(case["d"], case["k0"], case["k1"], case["j"], case["gws"]) == want


@pytest.mark.parametrize("row", golden106()["positions"], ids=lambda r: f"n{r['n']}")
def test_106bit_positions_equal_reference_kernels(orc, row):
    assert orc.ref_normal106(row["n"]) == (float.fromhex(row["hi"]), float.fromhex(row["lo"]))


def test_106bit_fixture_covers_the_cases():
    have = {(c["d"], c["k0"], c["k1"], c["j"], c["gws"]) for c in golden106()["cases"]}
    assert have == set(ref_dd106_cases())


def golden_od():
    with open(REF_OD_GOLDEN) as f:
        return json.load(f)["cases"]


def test_overdouble_fixture_covers_the_cases():
    have = {(c["d"], c["k0"], c["k1"], c["j"], c["precision"], c["gws"]) for c in golden_od()}
    assert have == set(ref_od_cases())


@pytest.mark.parametrize("case", golden_od(), ids=case_id)
def test_overdouble_replay_equals_reference_kernels(orc, case):
    """The -D OVERDOUBLE build's kernels (dd_div of split numerators and moduli) replayed."""
    d, k0, k1, j, p, gws = (case[k] for k in ("d", "k0", "k1", "j", "precision", "gws"))
    if p == 0:
        want = (float.fromhex(case["hi"]), float.fromhex(case["lo"]))
        assert orc.ref_job106(d, k0, k1, j, gws) == want
    else:
        got = (orc.ref_batch192(d, k0, k1, j, p, gws) << SHIFT[j]) & M192
        assert got == int(case["part"], 16)


@pytest.mark.parametrize("case", [c for c in ref_od_cases() if c[4] == 1],
                         ids=lambda c: "d{}_k{}-{}_j{}".format(*c[:4]))
def test_overdouble_replay_close_to_exact(orc, case):
    """The -D OVERDOUBLE kernels stay within the same 2 units of 2^-192 per term of the exact sum
    (the split numerator is exact, so dd_div adds no error of its own past 2^53)."""
    d, k0, k1, j, p, gws = case
    exact = py_batch_calc(d, k0, k1, j)
    diff = (orc.ref_batch192(d, k0, k1, j, p, gws) - exact) & M192
    diff = diff - (1 << 192) if diff >> 191 else diff
    terms = min(k1, d) - k0 + (20 if k1 >= d else 0)
    assert abs(diff) <= 2 * terms, diff

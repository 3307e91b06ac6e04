"""CPU tests of libclbellard.so: the library loads, exports every symbol include/clbellard.h
declares, and its host-side logic (normalisation, formula table, formats, work-file summer,
error codes) agrees with the oracle.  No term computation runs here (there is no GPU)."""
import ctypes
import os
import random
import re

import pytest

import clbellard_amd as clb
from clb_testlib import (DEN0, DEN1, NUMER, NUMESIGN, ROOT, i2w, py_normalize, w2i)

R = random.Random(7)
HEADER = os.path.join(ROOT, "include", "clbellard.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clb_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    handle = ctypes.CDLL(clb.LIB_PATH)
    for name in names:
        assert hasattr(handle, name), name
    # and the Python mirror binds every one of them
    assert set(names) <= set(clb._SIGNATURES)


def test_version():
    assert b"gfx950" in clb.lib.clb_version()


def test_normalize_matches_oracle(oracle):
    for n in [5, 6, 7, 8, 9, 10, 999, 1000, 10**12, 999999999999999999] + \
             [R.randrange(5, 10**18) for _ in range(3000)]:
        assert clb.normalize(n) == oracle.normalize(n)
    for bad in (0, 1, 4, 10**18, 2**64 - 1):
        with pytest.raises(clb.ClbError) as e:
            clb.normalize(bad)
        assert e.value.code == -1


def test_formula_table():
    for j in range(7):
        assert clb.formula(j) == (NUMER[j], NUMESIGN[j], DEN0[j], DEN1[j])
    with pytest.raises(clb.ClbError):
        clb.formula(7)


def test_hex48_matches_oracle(oracle):
    for _ in range(300):
        v = R.randrange(0, 2**192)
        assert clb.hex48(i2w(v)) == oracle.hex48(v) == f"{v:048x}"


def test_hex_dd_matches_oracle(oracle):
    cases = [(0.5, 0.0), (0.5, -2.0**-60), (0.5, -2.0**-80), (-0.5, -2.0**-60), (0.0, 0.0),
             (0.999999999, 1e-20), (-1e-30, 0.0)]
    for _ in range(2000):
        hi = R.random() * R.choice([1, -1])
        lo = hi * 2.0**-54 * (R.random() - 0.5)
        cases.append((hi, lo))
    for hi, lo in cases:
        assert clb.hex_dd(hi, lo) == oracle.hex_dd(hi, lo), (hi, lo)


def test_u192_to_dd_rounds_to_nearest():
    from fractions import Fraction
    for _ in range(2000):
        v = R.randrange(0, 2**192)
        hi, lo = clb.u192_to_dd(i2w(v))
        exact = Fraction(v, 2**192)
        assert hi == float(exact)  # Fraction -> float rounds to nearest even
        assert lo == float(exact - Fraction(hi))
    assert clb.u192_to_dd((0, 0, 0)) == (0.0, 0.0)


def _write_work(dirpath, j, b, hex48, crlf=True):
    os.makedirs(os.path.join(dirpath, str(j)), exist_ok=True)
    nl = "\r\n" if crlf else "\n"
    with open(os.path.join(dirpath, str(j), f"{j}_{b}.txt"), "w", newline="") as f:
        f.write(hex48 + nl + "0\tdevice" + nl + "12ms")


def test_work_sum_first_gap_and_mod(tmp_path, oracle):
    d = str(tmp_path)
    vals = {}
    acc = (ctypes.c_int * 48)()
    for j in range(7):
        for b in range(R.randrange(0, 4)):
            v = R.randrange(0, 2**192)
            vals[(j, b)] = v
            _write_work(d, j, b, f"{v:048x}", crlf=bool(b % 2))
            oracle.lib.orc_work_sum_add(acc, f"{v:048x}".encode())
    # a file after a gap is ignored (workmodeSum0_6.hsp:14-16)
    _write_work(d, 3, 9, "f" * 48)
    hexsum, counts = clb.work_sum(d)
    expect = sum(vals.values()) % 2**192
    assert hexsum == f"{expect:048x}"
    assert hexsum == "".join("0123456789abcdef"[x] for x in acc)
    assert counts == [sum(1 for (jj, _) in vals if jj == j) for j in range(7)]


def test_work_sum_rejects_malformed(tmp_path):
    _write_work(str(tmp_path), 0, 0, "ABCDEF" + "0" * 42)
    with pytest.raises(clb.ClbError) as e:
        clb.work_sum(str(tmp_path))
    assert e.value.code == -6
    _write_work(str(tmp_path), 0, 0, "12345")
    with pytest.raises(clb.ClbError):
        clb.work_sum(str(tmp_path))


def test_work_sum_empty_dir(tmp_path):
    assert clb.work_sum(str(tmp_path)) == ("0" * 48, [0] * 7)


def test_context_without_gpu_fails_loudly():
    n = ctypes.c_int()
    if clb.lib.clb_get_dev_count(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(clb.ClbError) as e:
        clb.Context(0)
    assert e.value.code == -2


def test_null_and_range_errors():
    ctx = ctypes.c_void_p()
    assert clb.lib.clb_ctx_create(0, 7, ctypes.byref(ctx)) == -1
    assert clb.lib.clb_hex48(None, None) == -1
    assert clb.lib.clb_batch_calc(None, 10, 0, 10, 0) == -1
    assert clb.lib.clb_last_error()

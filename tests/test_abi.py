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


def test_missing_library_fails_loudly(tmp_path):
    import subprocess, sys
    code = ("import clbellard_amd as c; c.LIB_PATH = %r; c._Lib._handle = None\n"
            "try:\n    c.normalize(10)\nexcept ImportError as e:\n    print('ImportError', e)\n")
    out = subprocess.run([sys.executable, "-c", code % str(tmp_path / "none.so")], cwd=ROOT,
                         capture_output=True, text=True)
    assert "ImportError" in out.stdout and "no CPU fallback" in out.stdout


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


def test_issue_rate_rejects_bad_arguments():
    out = (ctypes.c_double * 4)()
    assert clb.lib.clb_issue_rate(0, -1, 10.0, out) == clb.EINVAL
    assert clb.lib.clb_issue_rate(0, 12, 10.0, out) == clb.EINVAL
    assert clb.lib.clb_issue_rate(0, 0, 0.0, out) == clb.EINVAL
    assert clb.lib.clb_issue_rate(0, 0, 10.0, None) == clb.EINVAL


def test_null_and_range_errors():
    ctx = ctypes.c_void_p()
    assert clb.lib.clb_ctx_create(0, 7, ctypes.byref(ctx)) == -1
    assert clb.lib.clb_hex48(None, None) == -1
    assert clb.lib.clb_batch_calc(None, 10, 0, 10, 0) == -1
    assert clb.lib.clb_last_error()
    # launch knobs: no context is an argument error, before any device work
    assert clb.lib.clb_set_graph(None, 1) == -1
    assert clb.lib.clb_set_tiling(None, 16, 0) == -1
    assert clb.lib.clb_set_launch(None, 4) == -1
    assert clb.lib.clb_autotune(None, 10**6, 100.0, None, None) == -1
    assert b"context" in clb.lib.clb_last_error()


def _wl(tmp_path, text, name="wl.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_parse_worklist_reference_sample(tmp_path):
    # the format of work/worklist_sample.txt: digits=, precision= with a comment, j_b lines, LF
    p = _wl(tmp_path, "digits=2000000000\nprecision=1 (0=106bit , 1=192bit)\n0_0\n1_0\n2_0\n"
                      "3_0\n4_0\n5_0\n6_0")
    assert clb.parse_worklist(p) == (2000000000, 1, [(j, 0) for j in range(7)])


def test_parse_worklist_variants(tmp_path):
    # CR LF, spaces, a blank line, thousands of jobs, digits clamped to the reference maximum
    jobs = "\r\n".join(f"{j}_{b}" for b in range(500) for j in range(7))
    p = _wl(tmp_path, "digits = 1 000 000\r\nprecision = 2\r\n\r\n" + jobs + "\r\n")
    d, prec, js = clb.parse_worklist(p)
    assert (d, prec, len(js)) == (1000000, 2, 3500)
    assert js[7] == (0, 1)
    p = _wl(tmp_path, "digits=99999999999999999999\nprecision=0\n", "big.txt")
    assert clb.parse_worklist(p)[:2] == (999999999999999999, 0)
    p = _wl(tmp_path, "digits=100\nprecision=x\n0_0\n", "badprec.txt")
    assert clb.parse_worklist(p)[1] == -1
    p = _wl(tmp_path, "digits=100\nprecision=1\n9_0\n", "badjob.txt")
    with pytest.raises(clb.ClbError) as e:
        clb.parse_worklist(p)
    assert e.value.code == -6
    with pytest.raises(clb.ClbError) as e:
        clb.parse_worklist(str(tmp_path / "missing.txt"))
    assert e.value.code == -4


def test_cli_sum_and_usage(tmp_path):
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    v1, v2 = R.randrange(0, 2**192), R.randrange(0, 2**192)
    _write_work(str(tmp_path), 0, 0, f"{v1:048x}")
    _write_work(str(tmp_path), 5, 0, f"{v2:048x}")
    out = subprocess.run([cli, "sum", "--workdir", str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0
    assert f"{(v1 + v2) % 2**192:048x}" in out.stdout
    assert "0:1" in out.stdout and "5:1" in out.stdout and "3:0" in out.stdout
    assert subprocess.run([cli], capture_output=True).returncode == 2
    assert subprocess.run([cli, "bogus", "1"], capture_output=True).returncode == 2


def test_cli_batch_rejects_bad_positions(tmp_path):
    """`clbellard batch` reads every position before it touches a GPU: an unreadable file or a
    line that is not a decimal number fails with exit status 1 and a message."""
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    bad = tmp_path / "bad.txt"
    bad.write_text("1000000\n12x4\n")
    for args, stdin in (([str(bad)], None), ([str(tmp_path / "missing.txt")], None),
                        (["-"], "250000\n1e6\n"), (["-"], "-5\n"), (["-"], "+250000\n"),
                        (["-"], "18446744073709551616\n"), (["-"], "1" * 5000 + "\n")):
        out = subprocess.run([cli, "batch", *args], input=stdin, capture_output=True, text=True)
        assert out.returncode == 1 and "cannot read positions" in out.stderr, (args, out.stderr)


def test_cli_rejects_malformed_options():
    """Every option and the position of `clbellard normal` are parsed strictly before any GPU
    call: trailing characters, signs, overflow, out-of-range precision, unknown kernel names and
    missing values print the usage (exit status 2)."""
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    for args in (["normal", "12abc"], ["normal", "-5"], ["normal", "+1000000"], ["normal", ""],
                 ["normal", "18446744073709551616"], ["normal"],
                 ["normal", "1000000", "--precision", "3"],
                 ["normal", "1000000", "--precision", "1x"],
                 ["normal", "1000000", "--precision"],
                 ["normal", "1000000", "--device", "-1"],
                 ["normal", "1000000", "--kernel", "fast"],
                 ["normal", "1000000", "--replay-gws", "0"],
                 ["normal", "1000000", "--replay-gws", "2e24"],
                 ["normal", "1000000", "--replay-gws", "1000"],
                 ["normal", "1000000", "--replay-gws", str(2**24 + 256)],
                 ["normal", "4"], ["normal", "1000000000000000000"],
                 ["bogus", "1"], ["sum", "--replay-gws", "100"],
                 ["batch", "-", "--autotune", "0"], ["batch", "-", "--autotune", "5ms"],
                 ["work", "w.txt", "--batch-size", "-1"]):
        out = subprocess.run([cli, *args], input="", capture_output=True, text=True)
        assert out.returncode == 2 and "usage" in out.stderr, (args, out.returncode, out.stderr)


def test_cli_batch_checks_positions_before_the_gpu():
    """`clbellard batch` runs clb_normalize (host-only) on every position before it creates a
    context: a position outside [5, 999999999999999999] prints its message and the usage (exit
    status 2), also on a host without a GPU."""
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    for stdin in ("1000000\n4\n", "1000000000000000000\n", "0\n"):
        out = subprocess.run([cli, "batch", "-"], input=stdin, capture_output=True, text=True)
        assert out.returncode == 2 and "usage" in out.stderr, (stdin, out.stderr)
        assert "invalid digits" in out.stderr, out.stderr


def test_write_work_file_round_trip(tmp_path, oracle):
    d = str(tmp_path / "work")
    vals = [R.randrange(0, 2**192) for _ in range(5)]
    for b, v in enumerate(vals):
        clb.write_work_file(d, 4, b, i2w(v), clb.PREC_U192, "0\tdev", 1234)
    body = open(os.path.join(d, "4", "4_2.txt"), "rb").read()
    assert body == f"{vals[2]:048x}\r\n0\tdev\r\n1234ms".encode()
    assert clb.work_sum(d) == (f"{sum(vals) % 2**192:048x}", [0, 0, 0, 0, 5, 0, 0])
    clb.write_work_file(d, 6, 0, i2w(vals[0]), clb.PREC_DD106, "x", 1)
    line0 = open(os.path.join(d, "6", "6_0.txt"), "rb").read().split(b"\r\n")[0].decode()
    hi, lo = clb.u192_to_dd(i2w(vals[0]))
    assert line0 == oracle.hex_dd(hi, lo) and line0.endswith("0" * 20)
    with pytest.raises(clb.ClbError):
        clb.write_work_file(d, 7, 0, i2w(1))


SCHED = {"device": "AMD Instinct MI355X", "terms_per_lane": 4, "min_tiles_per_wave": 0,
         "blocks_per_cu": 6, "waves_per_group": 2, "min_chunk": 1024, "chunks_per_wave": 0,
         "max_replicas": 4}


def test_schedule_file_round_trip(tmp_path):
    p = str(tmp_path / "tune.txt")
    clb.write_schedule(p, SCHED)
    text = open(p).read()
    assert text.splitlines()[0] == "device=AMD Instinct MI355X" and "min_chunk=1024" in text
    assert clb.read_schedule(p) == SCHED
    # any key order, CR LF line ends and blank lines are read the same
    lines = text.splitlines()
    (tmp_path / "b.txt").write_bytes(("\r\n".join(lines[::-1] + [""]) + "\r\n").encode())
    assert clb.read_schedule(str(tmp_path / "b.txt")) == SCHED


def test_schedule_file_rejects_malformed(tmp_path):
    good = [f"{k}={v}" for k, v in SCHED.items()]
    cases = {
        "unknown key": (good + ["tile=4"], clb.EFORMAT),
        "repeated key": (good + ["min_chunk=64"], clb.EFORMAT),
        "repeated device": (good + ["device=x"], clb.EFORMAT),
        "missing key": ([x for x in good if not x.startswith("max_replicas")], clb.EFORMAT),
        "missing device": (good[1:], clb.EFORMAT),
        "no equals sign": (good + ["max_replicas"], clb.EFORMAT),
        "not an integer": ([x if not x.startswith("min_chunk") else "min_chunk=1e3"
                            for x in good], clb.EFORMAT),
        "blank value": ([x if not x.startswith("blocks_per_cu") else "blocks_per_cu="
                         for x in good], clb.EFORMAT),
        "huge value": ([x if not x.startswith("min_chunk") else "min_chunk=99999999999"
                        for x in good], clb.EFORMAT),
        "out of range": ([x if not x.startswith("blocks_per_cu") else "blocks_per_cu=65"
                          for x in good], clb.EINVAL),
        "bad tile length": ([x if not x.startswith("terms_per_lane") else "terms_per_lane=8"
                             for x in good], clb.EINVAL),
        "bad replicas": ([x if not x.startswith("max_replicas") else "max_replicas=6"
                          for x in good], clb.EINVAL),
        "empty device": (["device="] + good[1:], clb.EINVAL),
        "long device": (["device=" + "x" * 300] + good[1:], clb.EINVAL),
    }
    for name, (lines, code) in cases.items():
        p = tmp_path / "t.txt"
        p.write_text("\n".join(lines) + "\n")
        with pytest.raises(clb.ClbError) as e:
            clb.read_schedule(str(p))
        assert e.value.code == code, (name, e.value)
    (tmp_path / "big.txt").write_text("\n" * 5000)
    with pytest.raises(clb.ClbError) as e:
        clb.read_schedule(str(tmp_path / "big.txt"))
    assert e.value.code == clb.EFORMAT
    with pytest.raises(clb.ClbError) as e:
        clb.read_schedule(str(tmp_path / "missing.txt"))
    assert e.value.code == clb.EIO
    bad = dict(SCHED, device="a\tb")
    with pytest.raises(clb.ClbError) as e:
        clb.write_schedule(str(tmp_path / "w.txt"), bad)
    assert e.value.code == clb.EINVAL and not os.path.exists(tmp_path / "w.txt")


def test_cli_rejects_malformed_tune_cache(tmp_path):
    """A --tune-cache file that is there but malformed is reported before any GPU call."""
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    bad = tmp_path / "tune.txt"
    bad.write_text("device=x\nterms_per_lane=3\n")
    for args in (["normal", "1000000", "--tune-cache", str(bad)],
                 ["batch", "-", "--tune-cache", str(bad)],
                 ["normal", "1000000", "--tune-cache"], ["normal", "1000000", "--tune-cache", ""]):
        out = subprocess.run([cli, *args], input="1000000\n", capture_output=True, text=True)
        assert out.returncode == 2 and "usage" in out.stderr, (args, out.stderr)

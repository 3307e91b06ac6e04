"""CPU: the host library's parsing and formatting entry points (clbellard_amd/csrc/clb_format.cpp:
worklists, work files, the summer, digit positions, hex and double-double formats) and the oracle,
built with AddressSanitizer and UndefinedBehaviorSanitizer (g++/gcc, no HIP) and driven with
random, malformed and boundary inputs by tests/sanitize/*.  A memory error or undefined behaviour
fails the run (-fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _build(cmd, tmp_path):
    out = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr[-4000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_host_entry_points_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_fuzz")
    _build(["g++", "-std=c++17", *SAN, "-o", exe,
            os.path.join(ROOT, "tests", "sanitize", "host_fuzz.cpp"),
            os.path.join(ROOT, "clbellard_amd", "csrc", "clb_format.cpp")], tmp_path)
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    out = subprocess.run([exe, str(scratch), "3000"], capture_output=True, text=True, env=ENV,
                         timeout=300)
    assert out.returncode == 0, (out.stdout[-2000:], out.stderr[-4000:])
    assert "0 failed checks" in out.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_fuzz")
    _build(["gcc", "-std=gnu11", *SAN, "-ffp-contract=off", "-pthread", "-o", exe,
            os.path.join(ROOT, "tests", "sanitize", "oracle_fuzz.c"),
            os.path.join(ROOT, "oracle", "clb_oracle.c"), "-lm"], tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert out.returncode == 0, (out.stdout[-2000:], out.stderr[-4000:])
    assert "0 failed checks" in out.stdout

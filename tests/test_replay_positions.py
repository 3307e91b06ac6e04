"""GPU: the reference replay at the headline position N = 10^12 against the reference kernels' own
output there (tests/golden/ref_positions.json): all 48 hex digits of the 192-bit modes, and the
106-bit mode's dev_ans to the bit, double-double rounding included.  Precision 2 takes about
19 s, precision 1 about 25 s and precision 0 about 21 s; CLB_REPLAY_QUICK=1 keeps precision 2
only."""
import json
import os
import struct
import sys
import time

import pytest

import clbellard_amd as clb
from clb_testlib import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_positions.json")) as f:
    ROWS = json.load(f)["positions"]


@pytest.mark.parametrize("row", ROWS, ids=lambda r: f"n{r['n']}_p{r['precision']}")
def test_replay_equals_reference_output(row):
    if row["precision"] != 2 and os.environ.get("CLB_REPLAY_QUICK"):
        pytest.skip("CLB_REPLAY_QUICK: precision 2 only")
    print(f"replay N={row['n']} precision {row['precision']}", file=sys.stderr, flush=True)
    with clb.Context(0, row["precision"]) as ctx:
        ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        t0 = time.perf_counter()
        frac, hexs, _ = ctx.normal(row["n"])
        dt = time.perf_counter() - t0
    print(f"  {hexs} in {dt:.2f} s (reference kernels: {row['reference_seconds']} s)",
          file=sys.stderr, flush=True)
    assert hexs == row["hex"]
    if row["precision"] == 0:  # dev_ans itself: the bit patterns of (hi, lo)
        hi, lo = (struct.unpack("<d", struct.pack("<Q", w))[0] for w in frac[:2])
        assert (hi, lo) == (float.fromhex(row["hi"]), float.fromhex(row["lo"]))


@pytest.mark.parametrize("p", [0, 1, 2])
def test_cli_replay_prints_the_reference_digits(p):
    """`clbellard normal N --precision p --replay` at N = 10^6 prints the reference kernels'
    digits, as the restatement pinned against them (orc_ref_*) computes them."""
    import subprocess
    from clb_testlib import ROOT, load_oracle
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    orc = load_oracle()
    if p == 0:
        want = clb.hex_dd(*orc.ref_normal106(10**6))
    else:
        want = f"{orc.ref_normal192(10**6, p):048x}"
    out = subprocess.run([cli, "normal", "1000000", "--precision", str(p), "--replay"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines()[1] == want


@pytest.mark.parametrize("p,replay", [(1, False), (0, False), (1, True), (0, True)])
def test_cli_batch_equals_normal(tmp_path, p, replay):
    """`clbellard batch` over positions in one launch prints, per position and in input order,
    the N' and digits `clbellard normal` prints for that position alone (exact mode and replay)."""
    import subprocess
    from clb_testlib import ROOT
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    positions = [1000000, 250003, 10**7, 99999]
    path = tmp_path / "positions.txt"
    path.write_text("\n".join(str(n) for n in positions) + "\n\n")
    extra = ["--replay"] if replay else []
    out = subprocess.run([cli, "batch", str(path), "--precision", str(p), "--graph", *extra],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert len(lines) == len(positions)
    for n, line in zip(positions, lines):
        one = subprocess.run([cli, "normal", str(n), "--precision", str(p), *extra],
                             capture_output=True, text=True, timeout=120)
        assert one.returncode == 0, one.stderr
        digits_line, hexs = one.stdout.splitlines()[:2]
        assert line == f"{digits_line.split('=')[1]} {hexs}", (n, line, hexs)

#!/usr/bin/env python3
"""Adds the reference kernels' own output at whole positions to tests/golden/ref_positions.json:
Allloop (Bellard_192.hsp:264-296) on the reference's kernels compiled from its sources into
oracle/_ref/ (`make -C oracle ref`), with its 2^24 work-items per launch, on an MI355X: per formula
Set0mem, BatchCalc and AddAns (ref_job), the answers accumulated as AddAns does.  The 192-bit
modes give 48 hex digits; the 106-bit mode gives dev_ans (hi, lo) as float.hex() and
hexadecimal_dd's 28 digits + 20 zeros.  Existing rows for other (n, precision) pairs are kept.

    python tests/golden/make_ref_position_fixtures.py N PRECISION [PRECISION ...] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, ".."), os.path.join(HERE, "..", "..")]  # tests/, repo root
import clbellard_amd as clb  # noqa: E402
from clb_testlib import GOLDEN, RefKernels, py_normalize  # noqa: E402

PATH = os.path.join(GOLDEN, "ref_positions.json")


def twosum(a, b):  # Bellard.cl:1-7 (Python floats are IEEE doubles, round to nearest)
    x = a + b
    t = x - a
    return x, (a - (x - t)) + (b - t)


def add_ans106(ans, part):
    """dev_ans <- sumfix(addcab(dev_ans, part)) (Sum7's last step, Bellard.cl:1349-1352)."""
    x, y = twosum(ans[0], part[0])
    x, y = twosum(x, (y + ans[1]) + part[1])
    if x >= 1.0:  # sumfix, :491-501
        x, t = twosum(x, -1.0)
        x, y = twosum(x, t + y)
    return x, y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("n", type=int)
    ap.add_argument("precision", type=int, nargs="+")
    ap.add_argument("--out", default=PATH)
    args = ap.parse_args()
    with open(PATH) as f:
        doc = json.load(f)
    rows = {(r["n"], r["precision"]): r for r in doc["positions"]}
    _, d = py_normalize(args.n)
    ref = RefKernels()
    try:
        for p in args.precision:
            # Allloop formula by formula (Set0mem, BatchCalc, AddAns on the reference kernels), so
            # that progress shows; dev_ans accumulates the Sum7 outputs as AddAns does.
            t0 = time.perf_counter()
            acc = (0.0, 0.0) if p == 0 else 0
            for j in range(7):
                part = ref.job(d, 0, d, j, p)
                if p == 0:
                    acc = add_ans106(acc, part)
                else:
                    acc = (acc + (part[0] << 128 | part[1] << 64 | part[2])) % (1 << 192)
                print(f"  precision {p} formula {j}: {time.perf_counter() - t0:.1f} s", flush=True)
            dt = time.perf_counter() - t0
            w = acc if p == 0 else (acc >> 128, (acc >> 64) & (2**64 - 1), acc & (2**64 - 1))
            if p == 0:
                row = {"n": args.n, "precision": 0, "hi": w[0].hex(), "lo": w[1].hex(),
                       "hex": clb.hex_dd(*w), "reference_seconds": round(dt, 2)}
            else:
                row = {"n": args.n, "precision": p, "hex": clb.hex48(w),
                       "reference_seconds": round(dt, 2)}
            print(f"  {row['hex']} in {dt:.2f} s", flush=True)
            rows[(args.n, p)] = row
    finally:
        ref.close()
    doc["positions"] = [rows[k] for k in sorted(rows)]
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {len(rows)} rows to {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

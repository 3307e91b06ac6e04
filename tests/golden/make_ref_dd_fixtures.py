#!/usr/bin/env python3
"""Writes tests/golden/ref_dd192.json and tests/golden/ref_dd106.json: the reference's own kernels
(compiled from its sources into oracle/_ref/ by `make -C oracle ref`, run through the OpenCL
runtime on an MI355X) on every case of clb_testlib.ref_dd_cases() (precision 1 and 2) and
ref_dd106_cases() (precision 0), plus whole positions in precision 0.  Each job entry is one
Set0mem + BatchCalc + AddAns (KernelFuncs.as:15-74): the formula's partial after Sum7's update,
48 hex digits for the 192-bit modes, the double-double (hi, lo) as float.hex() strings for the
106-bit mode.  These are the reference's results, double-double rounding included, that the CPU
restatement (orc_ref_batch192, orc_ref_job106, orc_ref_normal106) is held to.

    python tests/golden/make_ref_dd_fixtures.py [out192.json out106.json]   (GPU and oracle/_ref)
"""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from clb_testlib import (REF_DD106_GOLDEN, REF_DD_GOLDEN, RefKernels, py_normalize,  # noqa: E402
                         ref_dd106_cases, ref_dd_cases, w2i)

SOURCE = "reference kernels Bellard.cl / Bellard_div2by1.cl, oracle/_ref, one MI355X"
POSITIONS106 = [10**6, 10**7]


def main(argv):
    out192 = argv[1] if len(argv) > 2 else REF_DD_GOLDEN
    out106 = argv[2] if len(argv) > 2 else REF_DD106_GOLDEN
    ref = RefKernels()
    rows, rows106, pos106 = [], [], []
    try:
        for d, k0, k1, j, p, gws in ref_dd_cases():
            ref.set_gws(gws)
            part = w2i(ref.job(d, k0, k1, j, p))
            rows.append({"d": d, "k0": k0, "k1": k1, "j": j, "precision": p, "gws": gws,
                         "part": f"{part:048x}"})
            print(f"d={d} k=[{k0},{k1}) j={j} p={p} gws={gws}: {part:048x}", flush=True)
        for d, k0, k1, j, gws in ref_dd106_cases():
            ref.set_gws(gws)
            hi, lo = ref.job(d, k0, k1, j, 0)
            rows106.append({"d": d, "k0": k0, "k1": k1, "j": j, "gws": gws,
                            "hi": hi.hex(), "lo": lo.hex()})
            print(f"106-bit d={d} k=[{k0},{k1}) j={j} gws={gws}: {hi!r} {lo!r}", flush=True)
        ref.set_gws()
        for n in POSITIONS106:
            hi, lo = ref.normal(py_normalize(n)[1], 0)
            pos106.append({"n": n, "hi": hi.hex(), "lo": lo.hex()})
            print(f"106-bit N={n}: {hi!r} {lo!r}", flush=True)
    finally:
        ref.set_gws()
        ref.close()
    with open(out192, "w") as f:
        json.dump({"source": SOURCE, "cases": rows}, f, indent=1)
    with open(out106, "w") as f:
        json.dump({"source": SOURCE, "cases": rows106, "positions": pos106}, f, indent=1)
    print(f"wrote {len(rows)} cases to {out192}, {len(rows106)} + {len(pos106)} to {out106}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))

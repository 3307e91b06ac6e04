#!/usr/bin/env python3
"""Writes tests/golden/ref_dd192.json: the reference's own precision-1/2 kernels (compiled from
its sources into oracle/_ref/ by `make -C oracle ref`, run through the OpenCL runtime on an
MI355X) on every case of clb_testlib.ref_dd_cases().  Each entry is one BatchCalc + AddAns
(KernelFuncs.as:25-74): the formula's partial after Sum7_192's shift, 48 hex digits.  These are
the reference's results, double-double rounding included, that the CPU restatement
(orc_ref_batch192) is held to.

    python tests/golden/make_ref_dd_fixtures.py [out.json]      (needs a GPU and oracle/_ref)
"""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from clb_testlib import REF_DD_GOLDEN, RefKernels, ref_dd_cases, w2i  # noqa: E402


def main(argv):
    out = argv[1] if len(argv) > 1 else REF_DD_GOLDEN
    ref = RefKernels()
    rows = []
    try:
        for d, k0, k1, j, p, gws in ref_dd_cases():
            ref.set_gws(gws)
            part = w2i(ref.job(d, k0, k1, j, p))
            rows.append({"d": d, "k0": k0, "k1": k1, "j": j, "precision": p, "gws": gws,
                         "part": f"{part:048x}"})
            print(f"d={d} k=[{k0},{k1}) j={j} p={p} gws={gws}: {part:048x}", flush=True)
    finally:
        ref.set_gws()
        ref.close()
    with open(out, "w") as f:
        json.dump({"source": "reference kernels Bellard.cl / Bellard_div2by1.cl, oracle/_ref, "
                             "one MI355X", "cases": rows}, f, indent=1)
    print(f"wrote {len(rows)} cases to {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))

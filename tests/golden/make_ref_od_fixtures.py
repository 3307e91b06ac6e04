#!/usr/bin/env python3
"""Writes tests/golden/ref_dd_overdouble.json: the reference's own kernels (Bellard.cl built with
-D OVERDOUBLE, as the reference host builds it once 10 d + 9 >= 2^53, Bellard_192.hsp:472-474;
Bellard_div2by1.cl's Refresh kernel for precision 2; all compiled from the reference sources into
oracle/_ref/ by `make -C oracle ref` and run through the OpenCL runtime on an MI355X) on every
job of clb_testlib.ref_od_cases().  Each entry is one Set0mem + BatchCalc + AddAns
(KernelFuncs.as:15-74): 48 hex digits for the 192-bit modes, the double-double (hi, lo) as
float.hex() strings for the 106-bit mode.  The CPU restatement (orc_ref_batch192, orc_ref_job106)
is held to these.

    python tests/golden/make_ref_od_fixtures.py [out.json]   (GPU and oracle/_ref)
"""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from clb_testlib import REF_OD_GOLDEN, RefKernels, ref_od_cases, w2i  # noqa: E402

SOURCE = ("reference kernels Bellard.cl (-D OVERDOUBLE) / Bellard_div2by1.cl, oracle/_ref, "
          "one MI355X")


def main(argv):
    out = argv[1] if len(argv) > 1 else REF_OD_GOLDEN
    ref = RefKernels()
    rows = []
    try:
        for d, k0, k1, j, p, gws in ref_od_cases():
            ref.set_gws(gws)
            row = {"d": d, "k0": k0, "k1": k1, "j": j, "precision": p, "gws": gws}
            if p == 0:
                hi, lo = ref.job(d, k0, k1, j, 0)
                row.update(hi=hi.hex(), lo=lo.hex())
            else:
                row["part"] = f"{w2i(ref.job(d, k0, k1, j, p)):048x}"
            rows.append(row)
            print(row, flush=True)
    finally:
        ref.set_gws()
        ref.close()
    with open(out, "w") as f:
        json.dump({"source": SOURCE, "cases": rows}, f, indent=1)
    print(f"wrote {len(rows)} cases to {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))

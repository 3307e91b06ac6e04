#!/usr/bin/env python3
"""Writes position_<N>.json: the exact-mode result at digit position N computed twice on an
MI355X, by the all-integer kernel (64-bit Montgomery, exact division) as the sum of term-range
jobs in several launches, and by the precision-split kernel (FP64 error-free modular squares) as
one normal-mode run; the two engines share no arithmetic, and the fixture is written only if their
48 hex digits agree.  The uneven-jobs workload at N = 2.5e12 (tests/test_workloads.py) is checked
against this value instead of spending another full run on the same position.

    python tests/golden/make_position_fixture.py 2500000000000 [output_dir]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import clbellard_amd as clb  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    n = int(sys.argv[1])
    out_dir = sys.argv[2] if len(sys.argv) > 2 else os.path.join(ROOT, "tests", "golden")
    n_out, d, _ = clb.normalize(n)
    chunks = 12  # launches of the all-integer kernel, each well under a minute at N = 2.5e12
    cuts = [d * i // chunks for i in range(chunks + 1)]
    total, t_plain = 0, 0.0
    with clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_PLAIN) as ctx:
        for c in range(chunks):
            end = cuts[c + 1] if c + 1 < chunks else d + 1  # the last chunk carries the tails
            t0 = time.perf_counter()
            parts = ctx.run_jobs([(d, cuts[c], end, j) for j in range(7)])
            t_plain += time.perf_counter() - t0
            total += sum(clb.u192_to_int(p) for p in parts)
            log(f"plain kernel: k in [{cuts[c]}, {min(end, d)}) done, {t_plain:.1f} s")
    plain = f"{total % (1 << 192):048x}"
    with clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_SPLIT) as ctx:
        t0 = time.perf_counter()
        _, split, _ = ctx.normal(n_out)
        t_split = time.perf_counter() - t0
    log(f"plain {plain} ({t_plain:.1f} s), split {split} ({t_split:.1f} s)")
    if plain != split:
        log("kernels disagree")
        return 1
    path = os.path.join(out_dir, f"position_{n_out}.json")
    with open(path, "w") as f:
        json.dump({"source": "tests/golden/make_position_fixture.py on one MI355X: the "
                             f"all-integer kernel over {chunks} launches of term-range jobs "
                             f"({t_plain:.1f} s) and the precision-split kernel's normal mode "
                             f"({t_split:.1f} s), equal digits",
                   "n": n_out, "hex": plain}, f, indent=1)
        f.write("\n")
    log(f"wrote {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

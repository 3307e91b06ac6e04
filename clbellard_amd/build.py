"""Build libclbellard.so (HIP kernels + C ABI host) and the clbellard CLI for gfx950.

    python -m clbellard_amd.build            # library + CLI
Outputs are written next to this file (in-tree) so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclbellard.so")
CLI = os.path.join(HERE, "clbellard")

ARCH = os.environ.get("CLB_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is written explicitly; the exact
# FP64 engine relies on the other products and sums being rounded on their own.
# -packed-fp32-ops off: keep the compiler from emitting packed FP32 instructions.
# Device code runs with the IEEE mode bit clear and FP64 denormals flushed: the hardware applies
# output modifiers (the mul:2 of the doubling square) only then.  Neither matters otherwise here:
# no value in the kernels is a NaN or a denormal (every one is an integer, 0, or 1/m for
# m < 2^64), and clang accepts the IEEE mode change only with NaNs declared absent.
BASE_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HIP_FLAGS = BASE_FLAGS + [
    "-Xarch_device", "-fdenormal-fp-math=preserve-sign",
    "-Xarch_device", "-mno-amdgpu-ieee", "-Xarch_device", "-fno-honor-nans",
]
# clb_replay.hip (the reference replay's double-double arithmetic) uses the same device mode: its
# values are finite and 0 or far above the denormal range, so rounding to nearest gives the same
# bits in either mode, and its exponent engine uses the doubling squares too.


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build_library() -> str:
    units = [("clb_kernels.hip", HIP_FLAGS), ("clb_host.cpp", HIP_FLAGS),
             ("clb_format.cpp", HIP_FLAGS),
             ("clb_replay.hip", HIP_FLAGS), ("clb_noalloc.hip", HIP_FLAGS),
             ("clb_peaks.hip", HIP_FLAGS)]
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory() as work:
        objs = []
        for src, flags in units:
            obj = os.path.join(work, os.path.splitext(src)[0] + ".o")
            _run([_hipcc(), *flags, "-c", "-o", obj, os.path.join(CSRC, src)])
            objs.append(obj)
        _run([_hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    src = os.path.join(CSRC, "clb_cli.cpp")
    tmp = CLI + ".tmp"
    _run([_hipcc(), "-O2", "-std=c++17", "-o", tmp, src, f"-L{HERE}", "-lclbellard",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle() -> str:
    """The CPU oracle used by tests and bench.py's cpu_baseline (not part of the product)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_reference_kernels(ref_dir: str = "/root/reference") -> bool:
    """The reference's OpenCL kernels compiled from their sources into oracle/_ref/, with the
    replay host (GPU-side reference for the tests; not part of the product).  Skipped, returning
    False, where the reference sources are absent."""
    if not os.path.exists(os.path.join(ref_dir, "Bellard.cl")):
        return False
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={ref_dir}"])
    return True


def main(argv=None) -> int:
    build_library()
    if os.path.exists(os.path.join(CSRC, "clb_cli.cpp")):
        build_cli()
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))

"""Writes tests/golden/pi_hex.json: hex digits of pi at chosen positions, from mpmath.

Run in the build container only (mpmath is a test-time dependency, never used by the product):
    python tests/golden/make_pi_fixtures.py

For each position N (1-based hex position after the point, a multiple of 5 as the reference's
digit normalisation produces) the fixture holds floor(frac(16^(N-1) * pi) * 2^192) as 48 hex
digits ("digits"), and floor(frac(16^(N-1) * pi) * 2^256) as 64 hex digits ("digits256", the
exact value against which the errors of 192-bit results are measured).  The exact value of the
reference's 192-bit sum differs from this by less than 362 * (d + 20) units of 2^-192 (d the loop
count), so these digits pin every bit above that.

"published" holds entries of D. H. Bailey's table ("The BBP Algorithm for Pi", 2006) of 14 hex
digits beginning at position N, as recalled (the table is not available offline).  The 10^6
entry is cross-checked against mpmath here; 10^7 .. 10^10 are checked by the GPU tests, where
they agree with both term kernels.
"""
import json
import os
import time

import mpmath

POSITIONS = [5, 10, 15, 20, 25, 30, 50, 100, 255 + 5, 500, 1000, 1005, 4100, 9995, 10000,
             65540, 100000, 123455, 250000, 500000, 777775, 1000000]

PUBLISHED = {
    "1000000": "26c65e52cb4593",
    "10000000": "17af5863efed8d",
    "100000000": "ecb840e21926ec",
    "1000000000": "85895585a0428b",
    "10000000000": "921c73c6838fb2",
    "1250000000000": "07e45733cc790b",
    "250000000000000": "e6216b069cb6c1",
}


def main() -> None:
    top = max(POSITIONS)
    t0 = time.time()
    mpmath.mp.prec = 4 * (top + 64) + 256
    pi = +mpmath.pi
    scale = mpmath.mpf(2) ** (4 * (top + 64))
    big = int(mpmath.floor(pi * scale))  # floor(pi * 2^(4 (top + 64)))
    total_bits = 4 * (top + 64)
    out, out256 = {}, {}
    for n in sorted(POSITIONS):
        assert n % 5 == 0
        # frac(16^(n-1) pi) * 2^192 = bits [4(n-1), 4(n-1) + 192) after the point
        shift = total_bits - 4 * (n - 1) - 192
        v = (big >> shift) & ((1 << 192) - 1)
        out[str(n)] = f"{v:048x}"
        v256 = (big >> (shift - 64)) & ((1 << 256) - 1)
        out256[str(n)] = f"{v256:064x}"
    assert out["1000000"].startswith(PUBLISHED["1000000"])
    data = {
        "source": "mpmath %s, pi to %d bits" % (mpmath.__version__, mpmath.mp.prec),
        "definition": "floor(frac(16^(N-1) * pi) * 2^192), 48 hex digits; digits256: "
                      "floor(frac(16^(N-1) * pi) * 2^256), 64 hex digits",
        "digits": out,
        "digits256": out256,
        "published": PUBLISHED,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pi_hex.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {path} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()

/* Sanitizer harness for the oracle (oracle/clb_oracle.c), built with gcc -fsanitize=address,undefined
 * by tests/test_sanitizers.py: random and boundary inputs for every function family, with the
 * consistency checks that need no reference (the 256-bit term's top bits are the 192-bit term,
 * range sums are partition invariant, normal() equals its 256-bit counterpart's top bits within
 * the floor losses).  Any memory error or undefined behaviour aborts the run. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../oracle/clb_oracle.h"

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static uint64_t next64(void) {
  rs += 0x9E3779B97F4A7C15ull;
  uint64_t z = rs;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main(void) {
  int bad = 0;
  for (int i = 0; i < 3000; ++i) {
    const int j = (int)(next64() % 7);
    uint64_t numer, den0, den1;
    int64_t sign;
    orc_formula(j, &numer, &sign, &den0, &den1);
    const uint64_t d = 1 + next64() % (i % 3 == 0 ? 1000ull : i % 3 == 1 ? 1000000000000ull
                                                                          : 100000000000000000ull);
    const uint64_t k = next64() % d;
    uint64_t q3[3], q4[4];
    orc_term192(d, k, sign, den0, den1, q3);
    orc_term256(d, k, sign, den0, den1, q4);
    if (q3[0] != q4[0] || q3[1] != q4[1] || q3[2] != q4[2]) ++bad;
    orc_tail192(d, sign, den0, den1, q3);
    orc_tail256(d, sign, den0, den1, q4);
    uint64_t nn, dd;
    int od;
    (void)orc_normalize(next64(), &nn, &dd, &od);
  }
  /* range sums: partition invariance on small ranges, 1 and 3 threads */
  for (int i = 0; i < 50; ++i) {
    const uint64_t d = 1 + (next64() % 5000) * 2;
    const int j = (int)(next64() % 7);
    const uint64_t k0 = next64() % d, k1 = k0 + next64() % (d - k0 + 1);
    const uint64_t cut = k0 + (k1 > k0 ? next64() % (k1 - k0 + 1) : 0);
    uint64_t a[3], b[3], c[3];
    orc_range_sum(d, k0, k1, j, 3, a);
    orc_range_sum(d, k0, cut, j, 1, b);
    orc_range_sum(d, cut, k1, j, 1, c);
    orc_add192(b, c);
    if (memcmp(a, b, sizeof a) != 0) ++bad;
    uint64_t x[4];
    orc_job256(d, k0, k1 + 5, j, 2, x);
    orc_batch_calc(d, k0, k1 + 5, j, 2, a);
  }
  /* whole positions: the 192-bit result is the 256-bit one's top bits up to the floor losses */
  const uint64_t ns[] = {5, 10, 100, 1005, 4100, 9995};
  for (size_t i = 0; i < sizeof ns / sizeof ns[0]; ++i) {
    uint64_t r3[3], r4[4];
    orc_normal(ns[i], 2, r3);
    orc_normal256(ns[i], 2, r4);
    const uint64_t diff = r4[2] - r3[2]; /* loss < 362 (d + 26) units of 2^-192 */
    if (!(r4[0] == r3[0] && diff < (1ull << 40))) ++bad;
    uint64_t ref[3];
    (void)orc_ref_normal192(ns[i], 1, 1 << 24, 2, ref);
    (void)orc_ref_normal192(ns[i], 2, 4096, 2, ref);
    char hex[49];
    orc_hex48(r3, hex);
  }
  double dd[2];
  (void)orc_ref_job106(399, 0, 400, 3, 1 << 24, 2, dd);
  (void)orc_ref_normal106(1005, 1 << 24, 2, dd);
  /* formats */
  for (int i = 0; i < 3000; ++i) {
    double v[2] = {ldexp((double)(next64() >> 11), -53) * ((next64() & 1) ? 1 : -1),
                   ldexp((double)(next64() >> 11), -53 - (int)(next64() % 80)) *
                       ((next64() & 1) ? 1 : -1)};
    char hex[49];
    orc_hex_dd(v, hex);
    int acc[48] = {0};
    char line[49];
    for (int c = 0; c < 48; ++c) line[c] = "0123456789abcdef"[next64() & 15];
    line[48] = 0;
    orc_work_sum_add(acc, line);
  }
  printf("oracle_fuzz: %d failed checks\n", bad);
  return bad ? 1 : 0;
}

/*
 * clb_oracle.c -- CPU restatement of CLBellard's exact-integer digit-extraction path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, the smoke check and bench.py's
 * cpu_baseline leg load this library; the product never does.  See clb_oracle.h for the
 * word order, the exact-integer semantics and the pinning status of each function.
 */
#include "clb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- host-side constants */

/* Bellard_192.hsp:449-450, :468-474. */
int orc_normalize(uint64_t n_in, uint64_t *n_out, uint64_t *d_out, int *overdouble) {
  if (n_in <= 4 || n_in > 999999999999999999ull) return -1;
  uint64_t d = ((n_in - 1) * 2 - 3) / 5;
  if (d % 2 == 0) d--;
  if (n_out) *n_out = (d * 5 + 3) / 2 + 1;
  if (d_out) *d_out = d;
  if (overdouble) *overdouble = (d * 10 + 9 >= 9007199254740992ull);
  return 0;
}

/* KernelFuncs.as:9-12 -- the seven terms of Bellard's formula. */
static const uint64_t kNumer[7] = {32, 1, 256, 64, 4, 4, 1};
static const int64_t kNumeSign[7] = {1, 1, 0, 1, 1, 1, 0};
static const uint64_t kDen0[7] = {4, 4, 10, 10, 10, 10, 10};
static const uint64_t kDen1[7] = {1, 3, 1, 3, 5, 7, 9};

int orc_formula(int j, uint64_t *numer, int64_t *numesign, uint64_t *den0, uint64_t *den1) {
  if (j < 0 || j > 6) return -1;
  if (numer) *numer = kNumer[j];
  if (numesign) *numesign = kNumeSign[j];
  if (den0) *den0 = kDen0[j];
  if (den1) *den1 = kDen1[j];
  return 0;
}

/* ---------------------------------------------------------------- Montgomery arithmetic */

/* Bellard.cl:211-215: Newton iteration x <- x(2 - m x) from x = m (3 correct bits), 4 steps. */
uint32_t orc_modinv32(uint32_t m) {
  uint32_t inv = m;
  for (int i = 0; i < 4; ++i) inv *= 2u - inv * m;
  return (uint32_t)0 - inv;
}

/* Bellard_div2by1.cl:255-262: same iteration on 64 bits, 6 steps. */
uint64_t orc_modinv64(uint64_t m) {
  uint64_t inv = m;
  for (int i = 0; i < 6; ++i) inv *= (uint64_t)2 - inv * m;
  return (uint64_t)0 - inv;
}

/* Bellard.cl:380-391.  x = xhi*2^32 + xlo < m*2^32; returns x * 2^-32 mod m in [0, m). */
uint32_t orc_mr32(uint32_t xlo, uint32_t xhi, uint32_t inv, uint32_t m) {
  uint32_t q = xlo * inv;
  uint32_t t = (uint32_t)(((uint64_t)q * m) >> 32);
  if (xlo != 0) t++; /* xlo + (q*m mod 2^32) is 0 or exactly 2^32 */
  uint32_t s = t + xhi;
  if (s >= m || s < xhi) s -= m;
  return s;
}

/* Bellard_div2by1.cl:264-274, the 64-bit analogue of orc_mr32. */
uint64_t orc_mr64(uint64_t xlo, uint64_t xhi, uint64_t inv, uint64_t m) {
  uint64_t q = xlo * inv;
  uint64_t t = (uint64_t)(((u128)q * m) >> 64);
  if (xlo != 0) t++;
  uint64_t s = t + xhi;
  if (s >= m || s < xhi) s -= m;
  return s;
}

static inline uint32_t mr32_mul(uint32_t a, uint32_t b, uint32_t inv, uint32_t m) {
  uint64_t p = (uint64_t)a * b;
  return orc_mr32((uint32_t)p, (uint32_t)(p >> 32), inv, m);
}

static inline uint64_t mr64_mul(uint64_t a, uint64_t b, uint64_t inv, uint64_t m) {
  u128 p = (u128)a * b;
  return orc_mr64((uint64_t)p, (uint64_t)(p >> 64), inv, m);
}

/* Bellard.cl:395-408: right-to-left binary powering of 1024 in Montgomery form. */
uint32_t orc_expmod32(uint64_t n, uint32_t m, uint32_t inv, uint32_t r2) {
  uint32_t p = orc_mr32(r2 << 10, r2 >> 22, inv, m); /* mont(1024) */
  uint32_t x = orc_mr32(r2, 0, inv, m);              /* mont(1) */
  for (int i = 0; i < 64; ++i) {
    if (n & 1) x = mr32_mul(x, p, inv, m);
    p = mr32_mul(p, p, inv, m);
    n >>= 1;
    if (n == 0) break;
  }
  return orc_mr32(x, 0, inv, m);
}

/* Bellard_div2by1.cl:276-305: left-to-right, fixed 3-bit windows, table b[i] = mont(1024^i). */
uint64_t orc_expmod64(uint64_t n, uint64_t m, uint64_t inv, uint64_t r2) {
  if (n == 0) return 1 % m;
  uint64_t b[8];
  b[0] = orc_mr64(r2, 0, inv, m);
  b[1] = orc_mr64(r2 << 10, r2 >> 54, inv, m);
  for (int i = 2; i < 8; ++i) b[i] = mr64_mul(b[i - 1], b[1], inv, m);
  int bits = 64 - __builtin_clzll(n);
  int first = bits % 3;
  if (first == 0) first = 3;
  int pos = bits - first;
  uint64_t x = b[(n >> pos) & ((1u << first) - 1)];
  while (pos > 0) {
    x = mr64_mul(x, x, inv, m);
    x = mr64_mul(x, x, inv, m);
    x = mr64_mul(x, x, inv, m);
    pos -= 3;
    unsigned digit = (unsigned)((n >> pos) & 7);
    if (digit) x = mr64_mul(x, b[digit], inv, m);
  }
  return orc_mr64(x, 0, inv, m);
}

/* ---------------------------------------------------------------- division by invariant integers */

/* Moller & Granlund, "Improved division by invariant integers" (2011), Algorithm 2, as used at
 * Bellard_div2by1.cl:105-156 with the table entry computed by division (CREATETABLEFLAG,
 * Bellard_div2by1.cl:1, :99-101).  d must have its top bit set; returns floor((2^128-1)/d) - 2^64. */
uint64_t orc_reciprocal_word(uint64_t d) {
  const uint64_t d0 = d & 1, d9 = d >> 55, d40 = (d >> 24) + 1, d63 = (d >> 1) + d0;
  const uint64_t v0 = ((1u << 19) - 3u * (1u << 8)) / d9;
  const uint64_t v1 = (v0 << 11) - ((v0 * v0 * d40) >> 40) - 1;
  const uint64_t v2 = (v1 << 13) + ((v1 * ((1ull << 60) - v1 * d40)) >> 47);
  const uint64_t e = (v2 >> 1) * d0 - v2 * d63;
  const uint64_t v3 = (v2 << 31) + ((uint64_t)(((u128)v2 * e) >> 64) >> 1);
  const u128 p = (u128)v3 * d;
  const uint64_t lo = (uint64_t)p, hi = (uint64_t)(p >> 64);
  return v3 - (hi + d + (lo + d < lo));
}

/* Algorithm 4 of the same paper (Bellard_div2by1.cl:162-177): (u1:u0) / d with u1 < d, d
 * normalized, v = orc_reciprocal_word(d).  Returns the quotient, *r the remainder. */
uint64_t orc_div2by1(uint64_t u1, uint64_t u0, uint64_t d, uint64_t *r, uint64_t v) {
  u128 q = (u128)v * u1 + (((u128)u1 << 64) | u0);
  uint64_t q1 = (uint64_t)(q >> 64) + 1, q0 = (uint64_t)q;
  uint64_t rem = u0 - q1 * d;
  if (rem > q0) {
    q1--;
    rem += d;
  }
  if (rem >= d) {
    q1++;
    rem -= d;
  }
  *r = rem;
  return q1;
}

static inline uint64_t shr64(uint64_t x, int n) { return n < 64 ? x >> n : 0; }

/* Bellard_div2by1.cl:195-211: schoolbook n-by-1 division with the normalized divisor. */
void orc_divnby1(int n, const uint64_t *u, uint64_t d, uint64_t *q) {
  const int l = __builtin_clzll(d);
  d <<= l;
  const uint64_t v = orc_reciprocal_word(d);
  uint64_t k = shr64(u[n - 1], 64 - l);
  for (int i = n - 1; i >= 1; --i) {
    uint64_t ui = (u[i] << l) | shr64(u[i - 1], 64 - l);
    q[i] = orc_div2by1(k, ui, d, &k, v);
  }
  q[0] = orc_div2by1(k, u[0] << l, d, &k, v);
}

/* Bellard_div2by1.cl:229-235: x*y mod n with n << s normalized and v its reciprocal. */
uint64_t orc_mulmodn(uint64_t x, uint64_t y, uint64_t n, int s, uint64_t v) {
  u128 p = (u128)x * y;
  uint64_t hi = (uint64_t)(p >> 64), lo = (uint64_t)p, r;
  orc_div2by1((hi << s) | shr64(lo, 64 - s), lo << s, n << s, &r, v);
  return r >> s;
}

/* ---------------------------------------------------------------- per-term evaluation */

uint64_t orc_pow1024_mod(uint64_t n, uint64_t m) {
  if (m < (1ull << 32)) {
    /* Bellard.cl:785-788: R = 2^32, r2 = 2^64 mod m, with the m == 1 guard of :787. */
    uint32_t inv = orc_modinv32((uint32_t)m);
    uint32_t r2 = (uint32_t)(UINT64_MAX % m) + 1;
    if (r2 == (uint32_t)m) r2 = 0;
    return orc_expmod32(n, (uint32_t)m, inv, r2);
  }
  /* Bellard_div2by1.cl:329-333: R = 2^64, r2 = R^2 mod m through the reciprocal. */
  int s = __builtin_clzll(m);
  uint64_t v = orc_reciprocal_word(m << s);
  uint64_t r1 = UINT64_MAX % m + 1;
  uint64_t r2 = orc_mulmodn(r1, r1, m, s, v);
  return orc_expmod64(n, m, orc_modinv64(m), r2);
}

/* Bellard_div2by1.cl:329-341.  Output in bigSum order (q[0] = most significant word). */
void orc_term192(uint64_t d, uint64_t k, int64_t numesign, uint64_t den0, uint64_t den1,
                 uint64_t q[3]) {
  const uint64_t m = den0 * k + den1;
  uint64_t r = orc_pow1024_mod(d - k, m);
  if ((numesign + (int64_t)(k % 2)) % 2 == 1) r = m - r; /* :335-337 */
  const uint64_t u[4] = {0, 0, 0, r};
  uint64_t w[4];
  orc_divnby1(4, u, m, w); /* w[3] = floor(r/m) is 0, or 1 when r == m (dropped mod 2^192) */
  q[0] = w[2];
  q[1] = w[1];
  q[2] = w[0];
}

void orc_add192(uint64_t acc[3], const uint64_t q[3]) {
  uint64_t s2 = acc[2] + q[2];
  uint64_t c = s2 < q[2];
  uint64_t s1 = acc[1] + q[1];
  uint64_t c1 = s1 < q[1];
  s1 += c;
  c1 |= (s1 < c);
  acc[0] = acc[0] + q[0] + c1;
  acc[1] = s1;
  acc[2] = s2;
}

/* Bellard.cl:1153-1253 computes the terms k = d+i, i = 0..19, whose value is
 * (-1)^(k + numesign) 2^(-10i) / m.  Each term here is floor(2^192 * frac(value)), the same
 * per-term rule the head uses: floor(2^(192-10i)/m), or its 192-bit complement when negative. */
void orc_tail192(uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1, uint64_t out[3]) {
  out[0] = out[1] = out[2] = 0;
  for (int i = 0; i < 20; ++i) {
    const uint64_t k = d + (uint64_t)i;
    const uint64_t m = den0 * k + den1;
    const int e = 192 - 10 * i;
    uint64_t u[4] = {0, 0, 0, 0};
    u[e / 64] = 1ull << (e % 64);
    uint64_t w[4];
    orc_divnby1(4, u, m, w);
    uint64_t q[3] = {w[2], w[1], w[0]};
    if ((numesign + (int64_t)(k % 2)) % 2 == 1) {
      q[0] = ~q[0];
      q[1] = ~q[1];
      q[2] = ~q[2];
    }
    orc_add192(out, q);
  }
}

/* Bellard.cl:1359-1398. */
void orc_sum7_192(uint64_t ans[3], uint64_t part[3], uint64_t numer) {
  int sht = 0;
  if (numer == 4) sht = 2;
  if (numer == 32) sht = 5;
  if (numer == 64) sht = 6;
  if (numer == 256) sht = 8;
  if (sht) {
    part[0] = (part[0] << sht) | (part[1] >> (64 - sht));
    part[1] = (part[1] << sht) | (part[2] >> (64 - sht));
    part[2] = part[2] << sht;
  }
  orc_add192(ans, part);
}

/* ---------------------------------------------------------------- 256-bit error oracle */

void orc_term256(uint64_t d, uint64_t k, int64_t numesign, uint64_t den0, uint64_t den1,
                 uint64_t q[4]) {
  const uint64_t m = den0 * k + den1;
  uint64_t r = orc_pow1024_mod(d - k, m);
  if ((numesign + (int64_t)(k % 2)) % 2 == 1) r = m - r; /* frac(-x) = (m - r) / m */
  const uint64_t u[5] = {0, 0, 0, 0, r};
  uint64_t w[5];
  orc_divnby1(5, u, m, w); /* w[4] = 1 only for r' = m, i.e. t = 0 (dropped mod 2^256) */
  q[0] = w[3];
  q[1] = w[2];
  q[2] = w[1];
  q[3] = w[0];
}

void orc_add256(uint64_t acc[4], const uint64_t q[4]) {
  unsigned __int128 c = 0;
  for (int i = 3; i >= 0; --i) {
    c += (unsigned __int128)acc[i] + q[i];
    acc[i] = (uint64_t)c;
    c >>= 64;
  }
}

void orc_tail256(uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1, uint64_t out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0;
  for (int i = 0; i <= 25; ++i) {
    const uint64_t k = d + (uint64_t)i;
    const uint64_t m = den0 * k + den1;
    const int e = 256 - 10 * i;
    uint64_t u[5] = {0, 0, 0, 0, 0};
    u[e / 64] = 1ull << (e % 64);
    uint64_t w[5];
    orc_divnby1(5, u, m, w);
    uint64_t q[4] = {w[3], w[2], w[1], w[0]};
    if ((numesign + (int64_t)(k % 2)) % 2 == 1) /* 2^256 - ceil(x 2^256), x 2^256 not an integer */
      for (int b = 0; b < 4; ++b) q[b] = ~q[b];
    orc_add256(out, q);
  }
}

struct range256_job {
  uint64_t d, k0, k1, step;
  int64_t numesign;
  uint64_t den0, den1;
  uint64_t acc[4];
};

static void *range256_worker(void *arg) {
  struct range256_job *jb = (struct range256_job *)arg;
  jb->acc[0] = jb->acc[1] = jb->acc[2] = jb->acc[3] = 0;
  for (uint64_t k = jb->k0; k < jb->k1; k += jb->step) {
    uint64_t q[4];
    orc_term256(jb->d, k, jb->numesign, jb->den0, jb->den1, q);
    orc_add256(jb->acc, q);
  }
  return NULL;
}

static void shl256(uint64_t v[4], int sh) {
  if (sh == 0) return;
  for (int i = 0; i < 3; ++i) v[i] = (v[i] << sh) | (v[i + 1] >> (64 - sh));
  v[3] <<= sh;
}

void orc_job256(uint64_t d, uint64_t k0, uint64_t k1, int j, int nthreads, uint64_t out[4]) {
  const uint64_t hi = k1 < d ? k1 : d;
  uint64_t offset = k0;
  out[0] = out[1] = out[2] = out[3] = 0;
  if (k0 < hi) {
    int nt = nthreads < 1 ? 1 : nthreads > 256 ? 256 : nthreads;
    if ((uint64_t)nt > hi - k0) nt = (int)(hi - k0);
    struct range256_job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < nt; ++t) {
      jobs[t].d = d;
      jobs[t].k0 = k0 + (uint64_t)t;
      jobs[t].k1 = hi;
      jobs[t].step = (uint64_t)nt;
      jobs[t].numesign = kNumeSign[j];
      jobs[t].den0 = kDen0[j];
      jobs[t].den1 = kDen1[j];
      if (nt == 1 || pthread_create(&th[t], NULL, range256_worker, &jobs[t]) != 0) {
        range256_worker(&jobs[t]);
        th[t] = 0;
      }
    }
    for (int t = 0; t < nt; ++t) {
      if (th[t]) pthread_join(th[t], NULL);
      orc_add256(out, jobs[t].acc);
    }
    offset = hi;
  }
  if (offset == d) {
    uint64_t t[4];
    orc_tail256(d, kNumeSign[j], kDen0[j], kDen1[j], t);
    orc_add256(out, t);
  }
  shl256(out, kNumer[j] == 1 ? 0 : __builtin_ctzll(kNumer[j]));
}

int orc_normal256(uint64_t n, int nthreads, uint64_t ans[4]) {
  uint64_t d;
  const int rc = orc_normalize(n, NULL, &d, NULL);
  ans[0] = ans[1] = ans[2] = ans[3] = 0;
  if (rc) return rc;
  for (int j = 0; j < 7; ++j) {
    uint64_t part[4];
    orc_job256(d, 0, d, j, nthreads, part);
    orc_add256(ans, part);
  }
  return 0;
}

/* ---------------------------------------------------------------- range sums */

struct range_job {
  uint64_t d, k0, k1, step;
  int64_t numesign;
  uint64_t den0, den1;
  uint64_t acc[3];
};

static void *range_worker(void *arg) {
  struct range_job *jb = (struct range_job *)arg;
  jb->acc[0] = jb->acc[1] = jb->acc[2] = 0;
  for (uint64_t k = jb->k0; k < jb->k1; k += jb->step) {
    uint64_t q[3];
    orc_term192(jb->d, k, jb->numesign, jb->den0, jb->den1, q);
    orc_add192(jb->acc, q);
  }
  return NULL;
}

void orc_range_sum(uint64_t d, uint64_t k0, uint64_t k1, int j, int nthreads, uint64_t out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (k1 <= k0) return;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if ((uint64_t)nthreads > k1 - k0) nthreads = (int)(k1 - k0);
  struct range_job jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].d = d;
    jobs[t].k0 = k0 + (uint64_t)t;
    jobs[t].k1 = k1;
    jobs[t].step = (uint64_t)nthreads;
    jobs[t].numesign = kNumeSign[j];
    jobs[t].den0 = kDen0[j];
    jobs[t].den1 = kDen1[j];
    if (nthreads == 1 || pthread_create(&th[t], NULL, range_worker, &jobs[t]) != 0) {
      range_worker(&jobs[t]);
      th[t] = 0;
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    if (th[t]) pthread_join(th[t], NULL);
    orc_add192(out, jobs[t].acc);
  }
}

/* KernelFuncs.as:25-66.  The launch chunking (GWS*16 / GWS*32 terms) does not change an exact
 * sum, so only the range boundaries and the tail condition (offset == d, :56-59) matter. */
void orc_batch_calc(uint64_t d, uint64_t k_start, uint64_t k_end, int j, int nthreads,
                    uint64_t out[3]) {
  const uint64_t hi = k_end < d ? k_end : d;
  uint64_t offset = k_start;
  out[0] = out[1] = out[2] = 0;
  if (k_start < hi) {
    orc_range_sum(d, k_start, hi, j, nthreads, out);
    offset = hi;
  }
  if (offset == d) {
    uint64_t t[3];
    orc_tail192(d, kNumeSign[j], kDen0[j], kDen1[j], t);
    orc_add192(out, t);
  }
}

/* Bellard_192.hsp:264-296. */
int orc_normal(uint64_t n, int nthreads, uint64_t ans[3]) {
  uint64_t d;
  int rc = orc_normalize(n, NULL, &d, NULL);
  ans[0] = ans[1] = ans[2] = 0;
  if (rc) return rc;
  for (int j = 0; j < 7; ++j) {
    uint64_t part[3];
    orc_batch_calc(d, 0, d, j, nthreads, part);
    orc_sum7_192(ans, part, kNumer[j]);
  }
  return 0;
}

/* ---------------------------------------------------------------- reference replay (dd) */

/* The reference's 192-bit kernels of precision 1 and 2 keep, per work-item, two double-double
 * sums (of r/m and of (r 2^96 mod m)/m) and convert them to 192 bits once; everything after is
 * exact integer addition.  The functions below restate that double-double path operation by
 * operation, including the fused multiply-adds that the reference's OpenCL build makes (OpenCL C
 * contracts a*b + c inside one expression; each is marked), so that the result is the
 * reference's to the bit.  The residues r are exact integers in the reference (Montgomery
 * exponentiation, Bellard.cl:395-461), so they are computed here with orc_pow1024_mod. */

typedef struct {
  double hi, lo;
} rdd;

/* Bellard.cl:1-7 */
static rdd rp_twosum(double a, double b) {
  rdd r;
  double x = a + b;
  double tmp = x - a;
  r.hi = x;
  r.lo = (a - (x - tmp)) + (b - tmp);
  return r;
}

/* Bellard.cl:9-15 (Veltkamp split by 2^27 + 1) */
static rdd rp_dsplit(double a) {
  rdd r;
  double tmp = a * 134217729.0;
  r.hi = tmp - (tmp - a);
  r.lo = a - r.hi;
  return r;
}

/* Bellard.cl:17-27; the error term's four products are fused (one expression). */
static rdd rp_twoproduct(double a, double b) {
  rdd r;
  r.hi = a * b;
  rdd ca = rp_dsplit(a), cb = rp_dsplit(b);
  double e = fma(ca.hi, cb.hi, -r.hi);
  e = fma(ca.lo, cb.hi, e);
  e = fma(ca.hi, cb.lo, e);
  r.lo = fma(ca.lo, cb.lo, e);
  return r;
}

/* Bellard.cl:29-35 */
static rdd rp_addcab(rdd a, rdd b) {
  rdd c = rp_twosum(a.hi, b.hi);
  c.lo = (c.lo + a.lo) + b.lo;
  return rp_twosum(c.hi, c.lo);
}

static rdd rp_neg(rdd a) {
  rdd r = {-a.hi, -a.lo};
  return r;
}

/* Bellard.cl:99-106 (nn_div) */
static rdd rp_nn_div(double x, double y) {
  double z1 = x / y;
  rdd c = rp_twoproduct(-z1, y);
  double z2 = ((c.hi + x) + c.lo) / y;
  return rp_twosum(z1, z2);
}

/* Bellard.cl:89-96 (nd_div); "- z1 * y.y" is fused with the sum before it. */
static rdd rp_nd_div(double x, rdd y) {
  double z1 = x / y.hi;
  rdd c = rp_twoproduct(-z1, y.hi);
  double z2 = (fma(-z1, y.lo, c.hi + x) + c.lo) / y.hi;
  return rp_twosum(z1, z2);
}

/* Bellard.cl:70-77 (dd_div); fused as in nd_div. */
static rdd rp_dd_div(rdd x, rdd y) {
  double z1 = x.hi / y.hi;
  rdd c = rp_twoproduct(-z1, y.hi);
  double z2 = ((fma(-z1, y.lo, c.hi + x.hi) + x.lo) + c.lo) / y.hi;
  return rp_twosum(z1, z2);
}

/* Bellard.cl:472-489 (sumfix2) */
static void rp_sumfix2(rdd *s) {
  double lo = 0.0;
  if (s->hi >= 1.0) lo = -1.0;
  if (s->hi <= -1.0) lo = 1.0;
  rdd t = rp_twosum(s->hi, lo);
  t.lo += s->lo;
  *s = rp_twosum(t.hi, t.lo);
}

/* Bellard.cl:506-578 (dd2_to_ulong3): two double-double sums -> 192 bits, w[0] most significant:
 * 64 bits from s0 (52 + 12 of the next 52), 40 + 24 more, then s1's 96 bits with the 2-bit
 * overlap check at bits 30-31 / 46-47. */
static void rp_dd2_to_ulong3(rdd s0, rdd s1, uint64_t w[3]) {
  const rdd one = {1.0, 0.0};
  if (s0.hi < 0.0) s0 = rp_addcab(s0, one);
  if (s1.hi < 0.0) s1 = rp_addcab(s1, one);
  uint64_t u0, u1, u2, ut;
  double t;
  s0.hi *= 4503599627370496.0; /* 2^52 */
  s0.lo *= 4503599627370496.0;
  t = trunc(s0.hi);
  u0 = (uint64_t)t * 4096u;
  s0.hi -= t;
  s0 = rp_twosum(s0.hi, s0.lo);
  s0.hi *= 4503599627370496.0;
  if (s0.hi < 0.0) {
    s0.hi += 4503599627370496.0;
    u0 -= 4096u;
  }
  ut = (uint64_t)s0.hi;
  u0 |= ut >> 40;
  u1 = (ut % (1ull << 40)) << 24;
  s1.hi *= 281474976710656.0; /* 2^48 */
  s1.lo *= 281474976710656.0;
  t = trunc(s1.hi);
  ut = (uint64_t)t;
  s1.hi -= t;
  s1 = rp_twosum(s1.hi, s1.lo);
  s1.hi *= 281474976710656.0;
  if (s1.hi < 0.0) {
    s1.hi += 281474976710656.0;
    ut--;
  }
  u2 = (ut % (1ull << 16)) << 48;
  u2 += (uint64_t)s1.hi;
  const uint32_t chk0 = (uint32_t)((u1 >> 30) % 4u);
  const uint32_t chk1 = (uint32_t)(ut >> 46);
  if (chk0 != chk1) {
    const uint64_t last = u1;
    if (((chk0 + 1u) % 4u) == chk1) {
      u1 += 1ull << 30;
      if (u1 < last) u0++;
    }
    if (((chk0 - 1u) % 4u) == chk1) {
      u1 -= 1ull << 30;
      if (u1 > last) u0--;
    }
  }
  u1 = (u1 & 0xFFFFFFFF00000000ull) | (ut >> 16);
  w[0] = u0;
  w[1] = u1;
  w[2] = u2;
}

/* The -D OVERDOUBLE build (Bellard_192.hsp:472-474): every kernel of the position's program is
 * built with it once 10 d + 9 >= 2^53. */
static int rp_overdouble(uint64_t d) { return d * 10 + 9 >= 9007199254740992ull; }

/* (double)v and the remainder, (double)((long)v - (long)(double)v) (Bellard.cl:938-939,
 * :1041-1042; v < 2^63). */
static rdd rp_u64dd(uint64_t v) {
  rdd r;
  r.hi = (double)v;
  r.lo = (double)((int64_t)v - (int64_t)r.hi);
  return r;
}

/* One term of Bellard.cl:783-805 / :1012-1069: the numerator is steered so that the sum stays
 * in (-1, 1) (m - r and a negated quotient while the sum is positive).  od: the -D OVERDOUBLE
 * build's dd_div of the split numerator and denominator (:957-962, :1040-1046) instead of
 * nn_div. */
static void rp_ddstep(rdd *s, uint64_t r, uint64_t m, int od) {
  const int pos = s->hi > 0.0;
  const uint64_t n2 = pos ? m - r : r;
  rdd q = od ? rp_dd_div(rp_u64dd(n2), rp_u64dd(m)) : rp_nn_div((double)n2, (double)m);
  if (pos) q = rp_neg(q);
  *s = rp_addcab(*s, q);
}

void orc_ref_item192(uint64_t d, uint64_t offset, uint64_t k_max, uint64_t idx, uint64_t gsize,
                     int j, uint64_t out[3]) {
  rdd s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
  uint64_t k = offset + idx;
  const int od = rp_overdouble(d);
  for (; k < k_max; k += gsize) {
    const uint64_t m = kDen0[j] * k + kDen1[j];
    const uint64_t r = orc_pow1024_mod(d - k, m);
    uint64_t b = (uint64_t)(((u128)r << 64) % m); /* r 2^96 mod m, Bellard.cl:798-799, :1052-1053 */
    b = (uint64_t)(((u128)b << 32) % m);
    /* the 32-bit kernel (m < 2^32) has no OVERDOUBLE variant */
    rp_ddstep(&s0, r, m, od && m >= (1ull << 32));
    rp_ddstep(&s1, b, m, od && m >= (1ull << 32));
  }
  if (((uint64_t)kNumeSign[j] + k % 2) % 2 == 1) { /* Bellard.cl:807-810, :1075-1078 */
    s0 = rp_neg(s0);
    s1 = rp_neg(s1);
  }
  rp_dd2_to_ulong3(s0, s1, out);
}

/* 2^e mod m for e <= 96 by doubling (the tail's numerators, exact in the reference). */
static uint64_t pow2_small(int e, uint64_t m) {
  uint64_t x = 1 % m;
  for (int i = 0; i < e; ++i) x = (uint64_t)(((u128)x << 1) % m);
  return x;
}

void orc_ref_tail192(uint64_t d, int j, uint64_t out[3]) {
  const uint64_t den0 = kDen0[j], den1 = kDen1[j];
  rdd s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
  double n0 = 1.0, n1 = 79228162514264337593543950336.0; /* 2^96 */
  for (int i = 0; i < 19; ++i) {
    n0 /= 1024.0;
    n1 /= 1024.0;
  }
  int64_t i = 19;
  for (; i >= 0; --i) { /* Bellard.cl:1171-1185 */
    const uint64_t m = den0 * (d + (uint64_t)i) + den1;
    const rdd y = {(double)m, (double)((int64_t)m - (int64_t)(double)m)};
    rdd q = rp_nd_div(n0, y);
    if ((d + (uint64_t)i) % 2 == 1) q = rp_neg(q);
    s0 = rp_addcab(s0, q);
    n0 *= 1024.0;
  }
  for (i = 19; i >= 10; --i) { /* :1187-1199 */
    const uint64_t m = den0 * (d + (uint64_t)i) + den1;
    const rdd y = {(double)m, (double)((int64_t)m - (int64_t)(double)m)};
    rdd q = rp_nd_div(n1, y);
    if ((d + (uint64_t)i) % 2 == 1) q = rp_neg(q);
    s1 = rp_addcab(s1, q);
    n1 *= 1024.0;
  }
  for (; i >= 0; --i) { /* :1202-1239: numerators 2^(96 - 10 i) mod m, exact */
    const uint64_t m = den0 * (d + (uint64_t)i) + den1;
    const uint64_t nmr = pow2_small(96 - 10 * (int)i, m);
    rdd q;
    if (m > 4294967296ull) {
      const rdd x = {(double)nmr, (double)((int64_t)nmr - (int64_t)(double)nmr)};
      const rdd y = {(double)m, (double)((int64_t)m - (int64_t)(double)m)};
      q = rp_dd_div(x, y);
    } else {
      q = rp_nn_div((double)nmr, (double)m);
    }
    if ((d + (uint64_t)i) % 2 == 1) q = rp_neg(q);
    s1 = rp_addcab(s1, q);
    rp_sumfix2(&s1);
  }
  if (kNumeSign[j] == 1) { /* :1241-1244 */
    s0 = rp_neg(s0);
    s1 = rp_neg(s1);
  }
  rp_dd2_to_ulong3(s0, s1, out);
}

struct item_job {
  uint64_t d, offset, k_max, gsize, i0, i1, step;
  int j;
  uint64_t acc[3];
};

static void *item_worker(void *arg) {
  struct item_job *jb = (struct item_job *)arg;
  jb->acc[0] = jb->acc[1] = jb->acc[2] = 0;
  for (uint64_t idx = jb->i0; idx < jb->i1; idx += jb->step) {
    uint64_t w[3];
    orc_ref_item192(jb->d, jb->offset, jb->k_max, idx, jb->gsize, jb->j, w);
    orc_add192(jb->acc, w);
  }
  return NULL;
}

/* One launch of Sglobal32mtg_192 / Sglobal64mtg_192 with gsize work-items: the sum of its
 * work-items' 192-bit values (the integer reductions that follow are exact). */
static void ref_launch192(uint64_t d, uint64_t offset, uint64_t k_max, uint64_t gsize, int j,
                          int nthreads, uint64_t out[3]) {
  const uint64_t items = k_max - offset < gsize ? k_max - offset : gsize;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if ((uint64_t)nthreads > items) nthreads = (int)items;
  struct item_job jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].d = d;
    jobs[t].offset = offset;
    jobs[t].k_max = k_max;
    jobs[t].gsize = gsize;
    jobs[t].i0 = (uint64_t)t;
    jobs[t].i1 = items;
    jobs[t].step = (uint64_t)nthreads;
    jobs[t].j = j;
    if (nthreads == 1 || pthread_create(&th[t], NULL, item_worker, &jobs[t]) != 0) {
      item_worker(&jobs[t]);
      th[t] = 0;
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    if (th[t]) pthread_join(th[t], NULL);
    orc_add192(out, jobs[t].acc);
  }
}

int orc_ref_batch192(uint64_t d, uint64_t k_start, uint64_t k_end, int j, int precision,
                     uint64_t gsize, int nthreads, uint64_t out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (j < 0 || j > 6 || (precision != 1 && precision != 2) || gsize == 0) return -1;
  if (d > (9223372036854775807ull - 9) / 10) return -1; /* moduli below 2^63, Bellard.cl:1041 */
  uint64_t k_max = k_start, offset = k_start;
  uint64_t lm32 = ((1ull << 32) - kDen1[j] - 1) / kDen0[j] + 1; /* KernelFuncs.as:31-33 */
  if (lm32 > d) lm32 = d;
  if (lm32 > k_end) lm32 = k_end;
  while (offset < lm32) { /* :34-41, GWS * 16 terms per launch */
    k_max = k_max + gsize * 16 < lm32 ? k_max + gsize * 16 : lm32;
    ref_launch192(d, offset, k_max, gsize, j, nthreads, out);
    offset = k_max;
  }
  const uint64_t lm64 = k_end < d ? k_end : d; /* :45-53, GWS * 32 terms per launch */
  if (precision == 2 && offset < lm64) {
    uint64_t part[3]; /* Sglobal64mtg_192_Refresh: exact per term, chunking irrelevant */
    orc_range_sum(d, offset, lm64, j, nthreads, part);
    orc_add192(out, part);
    offset = lm64;
  }
  while (offset < lm64) {
    k_max = k_max + gsize * 32 < lm64 ? k_max + gsize * 32 : lm64;
    ref_launch192(d, offset, k_max, gsize, j, nthreads, out);
    offset = k_max;
  }
  if (offset == d) { /* :56-59 */
    uint64_t t[3];
    orc_ref_tail192(d, j, t);
    orc_add192(out, t);
  }
  return 0;
}

int orc_ref_normal192(uint64_t n, int precision, uint64_t gsize, int nthreads, uint64_t ans[3]) {
  uint64_t d;
  ans[0] = ans[1] = ans[2] = 0;
  if (orc_normalize(n, NULL, &d, NULL)) return -1;
  for (int j = 0; j < 7; ++j) {
    uint64_t part[3];
    if (orc_ref_batch192(d, 0, d, j, precision, gsize, nthreads, part)) return -1;
    orc_sum7_192(ans, part, kNumer[j]);
  }
  return 0;
}

/* ---------------------------------------------------------------- reference replay, 106-bit mode */

/* The 106-bit mode (precision 0) keeps double-doubles all the way: per work-item sums, a 256-lane
 * tree per work-group (shared_reduction_dd), 65536 per-group accumulators updated launch by launch
 * (bigSum), the tail added into bigSum[0], two Ssum trees and Sum7.  Restated literally. */

#define REF_GROUPS 65536 /* BIGSUM_SIZE, Bellard_192.hsp:5 */

/* Bellard.cl:37-43 (subcab) */
static rdd rp_subcab(rdd a, rdd b) {
  rdd c = rp_twosum(a.hi, -b.hi);
  c.lo = (c.lo + a.lo) - b.lo;
  return rp_twosum(c.hi, c.lo);
}

/* Bellard.cl:491-501 (sumfix) */
static void rp_sumfix(rdd *s) {
  if (s->hi >= 1.0) {
    rdd t = rp_twosum(s->hi, -1.0);
    t.lo += s->lo;
    *s = rp_twosum(t.hi, t.lo);
  }
}

/* One work-item of Sglobal32mtg (Bellard.cl:729-752: sumfix after each term) or Sglobal64mtg
 * (:914-975: the steered numerator of the 192-bit kernels), with the sign of its terms. */
static rdd rp_item106(uint64_t d, uint64_t offset, uint64_t k_max, uint64_t idx, uint64_t gsize,
                      int j, int is64) {
  rdd s = {0.0, 0.0};
  uint64_t k = offset + idx;
  for (; k < k_max; k += gsize) {
    const uint64_t m = kDen0[j] * k + kDen1[j];
    const uint64_t r = orc_pow1024_mod(d - k, m);
    if (is64) {
      rp_ddstep(&s, r, m, rp_overdouble(d));
    } else {
      s = rp_addcab(s, rp_nn_div((double)r, (double)m));
      rp_sumfix(&s);
    }
  }
  if (((uint64_t)kNumeSign[j] + k % 2) % 2 == 1) s = rp_neg(s);
  return s;
}

/* shared_reduction_dd (Bellard.cl:583-597) and Ssum's tree (:1274-1297) over p[0..255]. */
static rdd rp_tree256(rdd *p) {
  for (int i = 128; i > 0; i /= 2)
    for (int l = 0; l < i; ++l) {
      rdd s = rp_addcab(p[l], p[l + i]);
      rp_sumfix2(&s);
      p[l] = s;
    }
  return p[0];
}

struct group_job {
  uint64_t d, offset, k_max, gsize;
  int j, is64;
  uint64_t g0, g1, step;
  rdd *bigsum;
};

static void *group_worker(void *arg) {
  struct group_job *jb = (struct group_job *)arg;
  rdd p[256];
  for (uint64_t g = jb->g0; g < jb->g1; g += jb->step) {
    for (int l = 0; l < 256; ++l)
      p[l] = rp_item106(jb->d, jb->offset, jb->k_max, g * 256 + (uint64_t)l, jb->gsize, jb->j,
                        jb->is64);
    rdd s = rp_addcab(jb->bigsum[g], rp_tree256(p)); /* :760-764, :979-983 */
    rp_sumfix2(&s);
    jb->bigsum[g] = s;
  }
  return NULL;
}

static void ref_launch106(uint64_t d, uint64_t offset, uint64_t k_max, uint64_t gsize, int j,
                          int is64, int nthreads, rdd *bigsum) {
  const uint64_t groups = gsize / 256;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if ((uint64_t)nthreads > groups) nthreads = (int)groups;
  struct group_job jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t] = (struct group_job){d, offset, k_max, gsize, j, is64, (uint64_t)t, groups,
                                 (uint64_t)nthreads, bigsum};
    if (nthreads == 1 || pthread_create(&th[t], NULL, group_worker, &jobs[t]) != 0) {
      group_worker(&jobs[t]);
      th[t] = 0;
    }
  }
  for (int t = 0; t < nthreads; ++t)
    if (th[t]) pthread_join(th[t], NULL);
}

/* Sglobal_after (Bellard.cl:1118-1150): k = d .. d+9, numerators 2^(-10 i). */
static rdd rp_tail106(uint64_t d, int j) {
  rdd s = {0.0, 0.0};
  double n = 1.0;
  for (uint64_t i = 0; i < 10; ++i) {
    const uint64_t m = kDen0[j] * (d + i) + kDen1[j];
    const rdd y = {(double)m, (double)((int64_t)m - (int64_t)(double)m)};
    rdd q = rp_nd_div(n, y);
    if ((d + i) % 2 == 1) q = rp_neg(q);
    s = rp_addcab(s, q);
    n /= 1024.0;
  }
  if (((uint64_t)kNumeSign[j] + 2) % 2 == 1) s = rp_neg(s);
  return s;
}

/* BatchCalc in the 106-bit mode on bigsum[REF_GROUPS] (updated as the reference's bigSumMem). */
static void ref_batch106(uint64_t d, uint64_t k_start, uint64_t k_end, int j, uint64_t gsize,
                         int nthreads, rdd *bigsum) {
  uint64_t k_max = k_start, offset = k_start;
  uint64_t lm32 = ((1ull << 32) - kDen1[j] - 1) / kDen0[j] + 1; /* KernelFuncs.as:31-33 */
  if (lm32 > d) lm32 = d;
  if (lm32 > k_end) lm32 = k_end;
  while (offset < lm32) { /* :34-41 */
    k_max = k_max + gsize * 16 < lm32 ? k_max + gsize * 16 : lm32;
    ref_launch106(d, offset, k_max, gsize, j, 0, nthreads, bigsum);
    offset = k_max;
  }
  const uint64_t lm64 = k_end < d ? k_end : d; /* :45-53 */
  while (offset < lm64) {
    k_max = k_max + gsize * 32 < lm64 ? k_max + gsize * 32 : lm64;
    ref_launch106(d, offset, k_max, gsize, j, 1, nthreads, bigsum);
    offset = k_max;
  }
  if (offset == d) bigsum[0] = rp_addcab(bigsum[0], rp_tail106(d, j)); /* :1149, no sumfix2 */
  rdd p[256], small[256];
  for (int b = 0; b < 256; ++b) { /* Ssum, 65536 -> 256 -> 1 (KernelFuncs.as:62-65) */
    memcpy(p, bigsum + 256 * b, sizeof p);
    small[b] = rp_tree256(p);
  }
  bigsum[0] = rp_tree256(small);
}

/* Sum7 (Bellard.cl:1324-1356): part <- frac(numer * part), ans <- sumfix(ans + part). */
static void rp_sum7_106(rdd *part, rdd *ans, uint64_t numer) {
  rdd a = *part;
  if (a.hi < 0.0) {
    const rdd one = {1.0, 0.0};
    a = rp_addcab(a, one);
  }
  a.hi *= (double)numer;
  a.lo *= (double)numer;
  const rdd t = {trunc(a.hi), 0.0};
  a = rp_subcab(a, t);
  *part = a;
  rdd s = rp_addcab(*ans, a);
  rp_sumfix(&s);
  *ans = s;
}

int orc_ref_job106(uint64_t d, uint64_t k_start, uint64_t k_end, int j, uint64_t gsize,
                   int nthreads, double out[2]) {
  if (j < 0 || j > 6 || gsize == 0 || gsize % 256 || gsize > 256ull * REF_GROUPS) return -1;
  if (d > (9223372036854775807ull - 9) / 10) return -1;
  rdd *bigsum = calloc(REF_GROUPS, sizeof(rdd));
  if (!bigsum) return -1;
  rdd ans = {0.0, 0.0};
  ref_batch106(d, k_start, k_end, j, gsize, nthreads, bigsum);
  rp_sum7_106(&bigsum[0], &ans, kNumer[j]);
  out[0] = bigsum[0].hi;
  out[1] = bigsum[0].lo;
  free(bigsum);
  return 0;
}

int orc_ref_normal106(uint64_t n, uint64_t gsize, int nthreads, double ans_out[2]) {
  uint64_t d;
  if (orc_normalize(n, NULL, &d, NULL)) return -1;
  if (gsize == 0 || gsize % 256 || gsize > 256ull * REF_GROUPS) return -1;
  rdd *bigsum = malloc(REF_GROUPS * sizeof(rdd));
  if (!bigsum) return -1;
  rdd ans = {0.0, 0.0};
  for (int j = 0; j < 7; ++j) { /* Bellard_192.hsp:264-296 */
    memset(bigsum, 0, REF_GROUPS * sizeof(rdd));
    ref_batch106(d, 0, d, j, gsize, nthreads, bigsum);
    rp_sum7_106(&bigsum[0], &ans, kNumer[j]);
  }
  ans_out[0] = ans.hi;
  ans_out[1] = ans.lo;
  free(bigsum);
  return 0;
}

/* ---------------------------------------------------------------- formats */

void orc_hex48(const uint64_t w[3], char out[49]) {
  static const char hx[] = "0123456789abcdef";
  for (int i = 0; i < 3; ++i)
    for (int b = 0; b < 16; ++b) out[16 * i + b] = hx[(w[i] >> (60 - 4 * b)) & 15];
  out[48] = 0;
}

static inline void dd_twosum(double a, double b, double *x, double *y) {
  double s = a + b;
  double t = s - a;
  *y = (a - (s - t)) + (b - t);
  *x = s;
}

/* hexsademical.as:12-19 (addcab): (a.0, a.1) + (b.0, b.1). */
static void dd_add(double *h, double *l, double bh, double bl) {
  double x, y;
  dd_twosum(*h, bh, &x, &y);
  y = y + *l + bl;
  dd_twosum(x, y, h, l);
}

/* hexsademical.as:21-101.  outdd is six 32-bit HSP integers viewed as 24 little-endian bytes;
 * bytes 23..10 receive the 14 emitted base-256 digits, with carries and borrows propagated
 * into the already emitted (more significant) bytes. */
/* HSP int() of a double (x86 cvttsd2si): INT_MIN for NaN and out-of-range values. */
static int hsp_int_of(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
  return (int)v;
}

void orc_hex_dd(const double dd[2], char out[49]) {
  double h = dd[0], l = dd[1];
  uint32_t outdd[6] = {0, 0, 0, 0, 0, 0};
  if (h < 0.0) dd_add(&h, &l, 1.0, 0.0);
  for (int cnt = 0; cnt < 14; ++cnt) {
    h *= 256.0;
    l *= 256.0;
    dd_twosum(h, l, &h, &l);
    int moji = hsp_int_of(h);
    if (moji < 256 && h >= 0.0) {
      ((uint8_t *)outdd)[23 - cnt] = (uint8_t)moji;
      h -= moji;
    } else {
      if (moji == 256) {
        int idx = 23 - cnt + 1;
        if (idx < 24) {
          int bidx = idx / 4;
          uint32_t addv = 1u << ((idx % 4) * 8);
          outdd[bidx] += addv;
          for (int rep = 5 - bidx; rep > 0; --rep) {
            if ((int32_t)outdd[bidx] >= 0 && (int32_t)outdd[bidx] < (int32_t)addv) {
              bidx++;
              addv = 1;
              outdd[bidx] += addv;
            } else {
              break;
            }
          }
        }
        ((uint8_t *)outdd)[23 - cnt] = 0;
        h -= 256.0;
      }
      if (h < 0.0) {
        moji = hsp_int_of(h + 256.0);
        int idx = 23 - cnt + 1;
        if (idx < 24) {
          int bidx = idx / 4;
          uint32_t addv = 1u << ((idx % 4) * 8);
          outdd[bidx] -= addv;
          for (int rep = 5 - bidx; rep > 0; --rep) {
            if ((int32_t)outdd[bidx] < 0 && (int32_t)outdd[bidx] >= -(int32_t)addv) {
              bidx++;
              addv = 1;
              outdd[bidx] -= addv;
            } else {
              break;
            }
          }
        }
        ((uint8_t *)outdd)[23 - cnt] = (uint8_t)moji;
        dd_add(&h, &l, 256.0, 0.0);
      }
    }
  }
  static const char hx[] = "0123456789abcdef";
  const uint8_t *bytes = (const uint8_t *)outdd;
  for (int i = 0; i < 24; ++i) {
    out[2 * i] = hx[bytes[23 - i] >> 4];
    out[2 * i + 1] = hx[bytes[23 - i] & 15];
  }
  out[48] = 0;
}

/* workmodeSum0_6.hsp:21-28: lowercase hex only ('a'..'f' map through "moji > 10: moji -= 39"),
 * single carry per digit, the carry out of digit 0 dropped (mod 16^48). */
void orc_work_sum_add(int acc[48], const char *line48) {
  for (int cnt = 0; cnt < 48; ++cnt) {
    int moji = (unsigned char)line48[47 - cnt] - 48;
    if (moji > 10) moji -= 39;
    acc[47 - cnt] += moji;
    if (acc[47 - cnt] >= 16) {
      acc[47 - cnt] -= 16;
      if (cnt != 47) acc[46 - cnt]++;
    }
  }
}

/*
 * clb_oracle.h -- CPU restatement of CLBellard's exact-integer digit-extraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The product
 * (clbellard_amd, libclbellard.so) never links, loads or calls it.
 *
 * Every function follows one piece of the reference and cites it (file:line).  192-bit values
 * are uint64_t[3] in the reference's bigSum order: w[0] = most significant word ("ulsum0"),
 * w[2] = least significant word ("ulsum2"), value = w[0]*2^128 + w[1]*2^64 + w[2], read as the
 * binary fraction value / 2^192.
 *
 * Exact-integer semantics.  The reference's "192bit(beta)" mode evaluates every head term with
 * m >= 2^32 as floor(r' * 2^192 / m) (Bellard_div2by1.cl:313-359), where r = 1024^(d-k) mod m
 * and r' = r for a positive term, m - r for a negative one.  This oracle applies that same rule
 * to every term: the head terms with m < 2^32 (where the reference keeps double-double sums,
 * Bellard.cl:771-823) and the 20 tail terms k = d..d+19 (Bellard.cl:1153-1253), whose value is
 * 2^(-10i)/m.  The result is independent of how the term range is partitioned.
 *
 * Pinning status (checked in tests/test_oracle.py):
 *   orc_normalize, orc_formula          closed form, reference presets
 *   orc_modinv32/64, orc_mr32/64,
 *   orc_expmod32/64, orc_pow1024_mod    brute force pow(2, 10n, m), random moduli < 2^62
 *   orc_reciprocal_word, orc_div2by1,
 *   orc_divnby1, orc_mulmodn            exact big-integer division
 *   orc_term192, orc_tail192            big-integer floor(r' 2^192 / m)
 *   orc_range_sum, orc_batch_calc       brute-force big-integer sums; partition invariance
 *   orc_sum7_192                        (c * S) mod 2^192
 *   orc_normal                          hex digits of pi from mpmath (tests/golden/)
 *   orc_term256, orc_tail256,
 *   orc_job256, orc_normal256           big-integer floor(r' 2^256 / m); floor(term256 / 2^64) ==
 *                                       term192; 256 bits of pi from mpmath at 22 positions
 *   orc_hex48                           printf("%016llx") layout
 *   orc_work_sum_add                    big-integer sum mod 16^48
 *   orc_hex_dd                          exact 112-bit values, hand-traced borrow vectors,
 *                                       floor(v 2^112) for random non-negative inputs, and for
 *                                       negative ones floor((x + s) 2^112), (x, y) = TwoSum(a0, 1),
 *                                       s = RN(y + a1): the add of 1's one rounding, computed with
 *                                       Python's IEEE sums and exact rationals
 *   orc_ref_item192, orc_ref_tail192,
 *   orc_ref_batch192, orc_ref_normal192,
 *   orc_ref_job106, orc_ref_normal106   the reference kernels' own results (GPU runs of
 *                                       oracle/_ref, committed as tests/golden/ref_dd192.json and
 *                                       ref_dd106.json): bit-exact; and within the double-double
 *                                       bound of the exact sums
 */
#ifndef CLB_ORACLE_H
#define CLB_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bellard_192.hsp:449-450 and :468-474.  Returns 0, or -1 for N <= 4 or N > 999999999999999999. */
int orc_normalize(uint64_t n_in, uint64_t *n_out, uint64_t *d_out, int *overdouble);

/* KernelFuncs.as:3-13: numerator c_j, sign flag (1 = minus), denominator den0*k + den1. */
int orc_formula(int j, uint64_t *numer, int64_t *numesign, uint64_t *den0, uint64_t *den1);

/* Bellard.cl:192-217 (modinv32) and Bellard_div2by1.cl:255-262 (modinv64_fast): -m^-1 mod 2^w. */
uint32_t orc_modinv32(uint32_t m);
uint64_t orc_modinv64(uint64_t m);

/* Bellard.cl:380-391 (MR32) and Bellard_div2by1.cl:264-274 (MR64_fast): Montgomery REDC. */
uint32_t orc_mr32(uint32_t xlo, uint32_t xhi, uint32_t inv, uint32_t m);
uint64_t orc_mr64(uint64_t xlo, uint64_t xhi, uint64_t inv, uint64_t m);

/* Bellard.cl:395-408 (ExpMod32v4): 1024^n mod m, R = 2^32, r2 = R^2 mod m. */
uint32_t orc_expmod32(uint64_t n, uint32_t m, uint32_t inv, uint32_t r2);
/* Bellard_div2by1.cl:276-305 (ExpMod64v4w3): 1024^n mod m, R = 2^64, 3-bit window. */
uint64_t orc_expmod64(uint64_t n, uint64_t m, uint64_t inv, uint64_t r2);

/* Moller-Granlund 2011, Alg. 2 / Alg. 4, as used at Bellard_div2by1.cl:105-177. */
uint64_t orc_reciprocal_word(uint64_t d);
uint64_t orc_div2by1(uint64_t u1, uint64_t u0, uint64_t d, uint64_t *r, uint64_t v);
/* Bellard_div2by1.cl:195-211: q[0..n-1] = floor(u / d), u and q least significant word first. */
void orc_divnby1(int n, const uint64_t *u, uint64_t d, uint64_t *q);
/* Bellard_div2by1.cl:229-235. */
uint64_t orc_mulmodn(uint64_t x, uint64_t y, uint64_t n, int s, uint64_t v);

/* 1024^n mod m with the reference's per-term setup: the 32-bit Montgomery path of
 * Bellard.cl:785-788 when m < 2^32, else Bellard_div2by1.cl:329-333. */
uint64_t orc_pow1024_mod(uint64_t n, uint64_t m);

/* One head term, Bellard_div2by1.cl:329-341: q = floor(r' * 2^192 / m) mod 2^192. */
void orc_term192(uint64_t d, uint64_t k, int64_t numesign, uint64_t den0, uint64_t den1,
                 uint64_t q[3]);
/* 192-bit add mod 2^192, Bellard_div2by1.cl:342-347 (with a plain carry chain). */
void orc_add192(uint64_t acc[3], const uint64_t q[3]);
/* Tail terms k = d..d+19, Bellard.cl:1153-1253, each as floor(2^192 * frac(+-2^(-10i)/m)). */
void orc_tail192(uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1, uint64_t out[3]);
/* Bellard.cl:1359-1398 (Sum7_192): part <<= log2(numer) mod 2^192 (written back), ans += part. */
void orc_sum7_192(uint64_t ans[3], uint64_t part[3], uint64_t numer);

/* Sum of head terms k in [k0, k1) for formula j (k1 <= d), split over nthreads host threads. */
void orc_range_sum(uint64_t d, uint64_t k0, uint64_t k1, int j, int nthreads, uint64_t out[3]);
/* KernelFuncs.as:25-66 (BatchCalc): head terms of [k_start, min(k_end, d)), plus the tail iff
 * the range reaches d; out = the formula's partial sum (before Sum7). */
void orc_batch_calc(uint64_t d, uint64_t k_start, uint64_t k_end, int j, int nthreads,
                    uint64_t out[3]);
/* Bellard_192.hsp:264-296 (Allloop): frac(16^(N'-1) * pi) as 192 bits. Returns orc_normalize's code. */
int orc_normal(uint64_t n, int nthreads, uint64_t ans[3]);

/* The error oracle (256-bit).  The exact value of a term is t = frac(+-2^(10(d-k)) / m) in [0, 1);
 * orc_term256 gives floor(t 2^256) (bigSum order, q[0] most significant) for a head term k < d,
 * from the same residue r = 1024^(d-k) mod m (orc_pow1024_mod) and a 5-word division.  Not the
 * reference's arithmetic: the reference truncates every term at 2^-192 (Bellard_div2by1.cl:338-341)
 * and the tail after 20 terms (Bellard.cl:1164); this carries 64 more bits per term and the tail to
 * k = d + 25 (the terms after that are below 2^-250 / m), so a sum of n terms here is below the
 * exact value by less than (n + 1) 2^-256, and the error of any 192-bit result is measured to that
 * accuracy. */
void orc_term256(uint64_t d, uint64_t k, int64_t numesign, uint64_t den0, uint64_t den1,
                 uint64_t q[4]);
/* 256-bit add mod 2^256 (q[0] most significant). */
void orc_add256(uint64_t acc[4], const uint64_t q[4]);
/* Tail terms k = d..d+25 of the formula (numesign, den0, den1), each floor(2^256 frac(+-2^(-10i)/m)). */
void orc_tail256(uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1, uint64_t out[4]);
/* Formula j's terms over [k0, min(k1, d)), plus the tail iff the range reaches d (BatchCalc's
 * rule, KernelFuncs.as:56-59), multiplied by c_j (mod 2^256): the 256-bit counterpart of one
 * clb_run_jobs result. */
void orc_job256(uint64_t d, uint64_t k0, uint64_t k1, int j, int nthreads, uint64_t out[4]);
/* frac(16^(N'-1) pi) to 256 bits (below the exact value by less than 362 (d + 27) 2^-256).
 * Returns orc_normalize's code. */
int orc_normal256(uint64_t n, int nthreads, uint64_t ans[4]);

/* Reference replay of the precision-1 and -2 kernels (bit-exact with the reference's own results,
 * double-double rounding included).  One work-item of Sglobal32mtg_192 (Bellard.cl:771-823) or
 * Sglobal64mtg_192 (:992-1084, the build without -D OVERDOUBLE): terms k = offset + idx + i gsize
 * < k_max, two double-double sums, dd2_to_ulong3 (:506-578). */
void orc_ref_item192(uint64_t d, uint64_t offset, uint64_t k_max, uint64_t idx, uint64_t gsize,
                     int j, uint64_t out[3]);
/* Sglobal_after_192 (Bellard.cl:1153-1253), formula j at loop count d. */
void orc_ref_tail192(uint64_t d, int j, uint64_t out[3]);
/* BatchCalc (KernelFuncs.as:25-66) with precision 1 (Sglobal32mtg_192 + Sglobal64mtg_192) or 2
 * (Sglobal32mtg_192 + the exact Refresh kernel), gsize work-items per launch (the reference
 * launches 2^24).  Returns -1 for a bad argument or 10 d + 9 >= 2^53 (the OVERDOUBLE build). */
int orc_ref_batch192(uint64_t d, uint64_t k_start, uint64_t k_end, int j, int precision,
                     uint64_t gsize, int nthreads, uint64_t out[3]);
/* Allloop (Bellard_192.hsp:264-296) on those kernels. */
int orc_ref_normal192(uint64_t n, int precision, uint64_t gsize, int nthreads, uint64_t ans[3]);

/* The 106-bit mode (precision 0) replayed: Set0memdd, BatchCalc with Sglobal32mtg / Sglobal64mtg
 * (Bellard.cl:729-766, :914-985), their 256-lane trees and the 65536 per-group accumulators,
 * Sglobal_after (:1118-1150), two Ssum trees (:1274-1297), then Sum7 (:1324-1356); out = bigSum[0]
 * after Sum7's update (what the reference's work mode saves), as (hi, lo).  gsize: work-items per
 * launch (a multiple of 256, at most 2^24). */
int orc_ref_job106(uint64_t d, uint64_t k_start, uint64_t k_end, int j, uint64_t gsize,
                   int nthreads, double out[2]);
/* Allloop in the 106-bit mode: dev_ans as (hi, lo). */
int orc_ref_normal106(uint64_t n, uint64_t gsize, int nthreads, double ans[2]);

/* Bellard_192.hsp:354-359 + hexsademical.as:104-127: 48 lowercase hex digits, w[0] first. */
void orc_hex48(const uint64_t w[3], char out[49]);
/* hexsademical.as:21-101 (hexadecimal_dd): double-double -> 28 hex digits + 20 zeros. */
void orc_hex_dd(const double dd[2], char out[49]);
/* workmodeSum0_6.hsp:21-28: acc (48 digit values 0..15, most significant first) += line. */
void orc_work_sum_add(int acc[48], const char *line48);

#ifdef __cplusplus
}
#endif
#endif

// clb_device.hpp -- gfx950 device arithmetic for the Bellard term kernels.
//
// Two exact engines compute, for a modulus m and exponent e, the 192-bit quotient
//   Q = floor(r * 2^192 / m) mod 2^192,   r = 2^e mod m,
// which is the per-term value of the reference's all-integer kernel (Bellard_div2by1.cl:329-341,
// with the sign applied later, see clb_kernels.hip).
//
//  * fp64 engine ("split"): every intermediate is an exact integer held in a double.  A modular
//    square uses the FMA error-free product (hi = x*y, lo = fma(x, y, -hi)), a quotient estimate
//    from a double reciprocal of m, and an exact FMA residue.  Residues stay in balanced form
//    |x| <= 1.02 m (up to 1.34 m for the widest moduli; fsq_bits gives the bounds), so no
//    correction step is needed inside the powering loop.  Valid while m * 2^c <= 2^50, c = the
//    window's largest shift (checked by the host).
//  * int engine ("plain"): 64-bit Montgomery (R = 2^64) built from 32x32->64 products, lazy
//    residues in [0, 2m), exponent windows folded into the REDC input, and the quotient words
//    obtained by exact division (Q word = s * (-m^-1) mod 2^64 for s = r 2^(64i) mod m).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clb {

// ------------------------------------------------------------------ 192-bit helpers (LSW first)

struct U192 {
  uint64_t w0, w1, w2;  // value = w2 * 2^128 + w1 * 2^64 + w0
};

__device__ __forceinline__ void add192(U192& a, uint64_t b0, uint64_t b1, uint64_t b2) {
  uint64_t s0 = a.w0 + b0;
  uint64_t c0 = s0 < b0;
  uint64_t s1 = a.w1 + b1;
  uint64_t c1 = s1 < b1;
  s1 += c0;
  c1 += s1 < c0;
  a.w2 = a.w2 + b2 + c1;
  a.w1 = s1;
  a.w0 = s0;
}

// a = -a - n  (mod 2^192): the sum of n negative terms -Q - 1 given a = sum of their Q.
__device__ __forceinline__ void neg192_minus(U192& a, uint64_t n) {
  // ~a = -a - 1, so -a - n = ~a - (n - 1) = ~a + (1 - n)
  U192 t = {~a.w0, ~a.w1, ~a.w2};
  uint64_t add = 1 - n;                   // low word of the signed value 1 - n
  uint64_t ext = (n > 1) ? ~0ull : 0;     // its sign extension: 1 - n < 0 iff n >= 2
  add192(t, add, ext, ext);
  a = t;
}

// ------------------------------------------------------------------ fp64 engine

// Exact integer in [0, 2^53] -> uint64 (v_floor + v_fma + two v_cvt_u32).
__device__ __forceinline__ uint64_t f2u64(double f) {
  double h = __builtin_floor(f * 0x1p-32);
  double l = __builtin_fma(h, -0x1p32, f);
  return ((uint64_t)(uint32_t)h << 32) | (uint64_t)(uint32_t)l;
}

// The two's complement of an exact integer |f| < 2^53 that may be negative.
__device__ __forceinline__ uint64_t f2i64(double f) {
  double h = __builtin_floor(f * 0x1p-32);
  double l = __builtin_fma(h, -0x1p32, f);
  return ((uint64_t)(int64_t)(int32_t)h << 32) | (uint64_t)(uint32_t)l;
}

// Reciprocal of m (an exact integer >= 1) to within about one ulp: hardware estimate plus two
// Newton steps.
__device__ __forceinline__ double recip(double m) {
  double r = __builtin_amdgcn_rcp(m);
  double e = __builtin_fma(-m, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-m, r, 1.0);
  return __builtin_fma(r, e, r);
}

// x * (x * 2^c) mod m in balanced form (|x| and the result within the bounds given at fsq_bits).
// Six FP64 operations plus the scaling: hi + lo is the exact product, q its rounded quotient, t
// the exact residue of hi.
__device__ __forceinline__ double fsq(double x, double m, double minv) {
  double hi = x * x;
  double lo = __builtin_fma(x, x, -hi);
  double q = __builtin_rint(hi * minv);
  double t = __builtin_fma(-q, m, hi);
  return t + lo;
}

__device__ __forceinline__ double fsq_shift(double x, int c, double m, double minv) {
  double y = __builtin_amdgcn_ldexp(x, c);
  double hi = x * y;
  double lo = __builtin_fma(x, y, -hi);
  double q = __builtin_rint(hi * minv);
  double t = __builtin_fma(-q, m, hi);
  return t + lo;
}

// One 48-bit quotient digit of the long division rem * 2^192 / m.  In: rem in [0, 2m) (or the
// balanced residue for the first digit).  Out: digit q with N - q m in [0, 2m), N = rem * 2^48.
// The estimate floor(N/m - 1/4 + err), |err| < 0.16, is floor(N/m) or one less.
__device__ __forceinline__ double fdigit(double& rem, double m, double minv) {
  double n = __builtin_amdgcn_ldexp(rem, 48);
  double q = __builtin_floor(__builtin_fma(n, minv, -0.25));
  rem = __builtin_fma(-q, m, n);
  return q;
}

// 2^e mod m (balanced) for an exponent whose bits above the lowest s form t = e >> s < 32, the
// s low bits consumed in W-bit windows, most significant first.  s is a multiple of W, uniform
// per wave.
template <int W>
__device__ __forceinline__ double fpow2(uint64_t e, int s, double m, double minv) {
  const uint32_t t = (uint32_t)(e >> s);
  const double p = __builtin_amdgcn_ldexp(1.0, (int)t);
  double x = __builtin_fma(-__builtin_rint(p * minv), m, p);
  for (int pos = s - W; pos >= 0; pos -= W) {
#pragma unroll
    for (int i = 0; i < W - 1; ++i) x = fsq(x, m, minv);
    const int c = (int)((e >> pos) & ((1u << W) - 1));
    x = fsq_shift(x, c, m, minv);
  }
  return x;
}

// x <- 2^(H's bits pos..1) applied by squaring: for pos = npos..1, x <- x^2 mod m, or 2 x^2 mod m
// where bit pos of H is set.  fsq's six instructions; the doubling is the output modifier
// (mul:2) of the final add, so it costs nothing, and bit pos is a scalar test and branch (H and
// pos are wave-uniform).  The hardware ignores output modifiers while the IEEE mode bit is set, so
// the build clears it for the device code, and flushes FP64 denormals as well, the mode in which
// the compiler uses output modifiers itself (no value here is a NaN or a denormal: every one is an
// integer, 0, or 1/m for m < 2^64).
// Bounds, with |x| <= y m and the estimate of x^2 / m off by at most y^2 m 1.5 2^-52: for
// m < 2^42, 2^46, 2^48 the result stays within y = 1.02, 1.06, 1.34 m, and fsq_shift (m 2^c <=
// 2^50.25) on such an input returns at most 0.97, 1.0, 1.3 m; every intermediate is an exact
// integer below 2^53.  The compiler folds no multiplication by 2.0 into an FP64 output modifier,
// and a branch between two squares made it copy x between registers on every bit, hence one
// block of assembly for the whole loop: ALU instructions and branches only, x kept in one
// register, every output used, all paths leaving at the end.  Needs npos >= 1.
__device__ __forceinline__ double fsq_bits(double x, uint64_t H, uint32_t npos, double m,
                                           double minv) {
  double hi, q;  // x's register takes lo, then the result
  asm volatile(
      "1:\n\t"
      "s_bitcmp1_b64 %[H], %[pos]\n\t"
      "v_mul_f64 %[hi], %[x], %[x]\n\t"
      "v_fma_f64 %[x], %[x], %[x], -%[hi]\n\t"
      "v_mul_f64 %[q], %[hi], %[minv]\n\t"
      "v_rndne_f64 %[q], %[q]\n\t"
      "v_fma_f64 %[q], -%[q], %[m], %[hi]\n\t"
      "s_cbranch_scc1 2f\n\t"
      "v_add_f64 %[x], %[q], %[x]\n\t"
      "s_branch 3f\n"
      "2:\n\t"
      "v_add_f64 %[x], %[q], %[x] mul:2\n"
      "3:\n\t"
      "s_sub_u32 %[pos], %[pos], 1\n\t"
      "s_cmp_lg_u32 %[pos], 0\n\t"
      "s_cbranch_scc1 1b"
      : [x] "+v"(x), [hi] "=&v"(hi), [q] "=&v"(q), [pos] "+s"(npos)
      : [H] "s"(H), [m] "v"(m), [minv] "v"(minv)
      : "scc");
  return x;
}

// 2^(1024 H + v) mod m (balanced) for a wave-uniform H >= 1 and a per-lane even v in [0, 2048):
// the lanes of a wave have exponents E - 10 L, which this splits into a common high part H and a
// per-lane low part v.  The leading bits t of H (everything above the lowest s bits, s the
// smallest count that leaves t <= tmax) give the start value 2^t - rint(2^t / m) m in one exact
// step: with 2^t / m <= 2^49 the quotient is off by at most 0.5 + 1.5 2^-3, so |x| <= 0.69 m;
// tmax (uniform) = 49 + floor(log2 m) of the smallest m the caller passes.  H's other bits are
// scalar, so each takes a plain or a doubling square by a scalar branch (fsq_bits), no ldexp;
// H's last bit also absorbs bit 10 of v (a per-lane shift of at most 2); then W-bit windows take
// v's low ten bits.  Valid while m 2^(2^W) <= 2^50.
template <int W>
__device__ __forceinline__ double fpow2_split(uint64_t H, uint32_t v, int tmax, double m,
                                              double minv) {
  constexpr uint32_t mask = (1u << W) - 1;
  // 32-bit compares only: gfx9 has no scalar 64-bit ordered compare, so a 64-bit one would cost
  // VALU instructions in every term.
  const int bl = 64 - __builtin_clzll(H);
  const int bt = 32 - __builtin_clz((uint32_t)tmax);
  int s = 0;
  if (bl > bt || (bl == bt && (uint32_t)H >= (uint32_t)tmax)) {  // s = 0 needs H + 1 <= tmax
    s = bl > bt ? bl - bt : 1;                                   // now H >> s < 2^bt
    if ((uint32_t)(H >> s) > (uint32_t)tmax) s += 1;
  }
  const uint32_t t = (uint32_t)(H >> s);
  const int carry = (int)(v >> 10);
  double x;
  if (s == 0) {
    const double p = __builtin_amdgcn_ldexp(1.0, (int)t + carry);
    x = __builtin_fma(-__builtin_rint(p * minv), m, p);
  } else {
    const double p = __builtin_bit_cast(double, (uint64_t)(1023 + t) << 52);  // 2^t, t uniform
    x = __builtin_fma(-__builtin_rint(p * minv), m, p);
    if (s > 1) x = fsq_bits(x, H, (uint32_t)(s - 1), m, minv);
    x = fsq_shift(x, (int)(H & 1) + carry, m, minv);
  }
  // v is even (every exponent 10 (d - k) is), so where bit 0 forms a window of its own (W = 3:
  // windows 7-9, 4-6, 1-3, then bit 0; W = 1) the last square needs no shift.
  constexpr int last = W == 2 ? 0 : 1;  // lowest position of a shifted window
#pragma unroll
  for (int pos = 10 - W; pos >= last; pos -= W) {
#pragma unroll
    for (int i = 0; i < W - 1; ++i) x = fsq(x, m, minv);
    x = fsq_shift(x, (int)((v >> pos) & mask), m, minv);
  }
  if constexpr (last == 1) x = fsq(x, m, minv);
  return x;
}

// ------------------------------------------------------------------ int engine

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// -m^-1 mod 2^64 for odd m: 5-bit seed (3m xor 2), three 32-bit Newton steps, one 64-bit step.
__device__ __forceinline__ uint64_t neg_inv64(uint64_t m) {
  const uint32_t m32 = (uint32_t)m;
  uint32_t x = (3u * m32) ^ 2u;
  x *= 2u - m32 * x;
  x *= 2u - m32 * x;
  x *= 2u - m32 * x;
  uint64_t X = x;
  X *= 2ull - m * X;
  return 0ull - X;
}

// REDC((hi:lo)): (hi:lo + q m) / 2^64 with q = lo * mp mod 2^64.  Input < m * 2^64.
__device__ __forceinline__ uint64_t redc(uint64_t hi, uint64_t lo, uint64_t m, uint64_t mp) {
  const uint64_t q = lo * mp;
  return hi + mulhi64(q, m) + (lo != 0);
}

// Montgomery square with the window shift folded in: REDC(h^2 * 2^c).  h < 2m (lazy), c <= 7,
// m < 2^55 keeps h^2 2^c < m 2^64, so the result is again < 2m.
__device__ __forceinline__ uint64_t isq_shift(uint64_t h, int c, uint64_t m, uint64_t mp) {
  uint64_t lo = h * h;
  uint64_t hi = mulhi64(h, h);
  hi = (hi << c) | ((lo >> 1) >> (63 - c));  // branch-free for c = 0
  lo <<= c;
  return redc(hi, lo, m, mp);
}

__device__ __forceinline__ uint64_t isq(uint64_t h, uint64_t m, uint64_t mp) {
  return redc(mulhi64(h, h), h * h, m, mp);
}

// The same two squares for m < 2^55 and lazy h < 2m, written as chained 32x32+64 multiply-adds
// (v_mad_u64_u32) on 32-bit words: h < 2^56 has an upper word below 2^24 and m one below 2^23, so
// no column carries out of 64 bits and the REDC needs no 128-bit intermediate.  The product
// T = h (h 2^c) (four words t0..t3, c <= 7) takes 4 products, 3 for the plain square
// (h0 (2 h1) folds the doubled cross term); q = (t1:t0) mp mod 2^64 three; q m + (t1:t0) four,
// with the low words t0, t1 as addends of the first two columns: that sum is a multiple of 2^64
// (its low words come out 0), so its high half is exactly (q m + (t1:t0)) / 2^64 and
// REDC(T) = (t3:t2) + that half, in [0, 2m), with no separate carry.  Every column fits 64 bits:
// (2^32 - 1)^2 + 2 (2^32 - 1) = 2^64 - 1.  Ten or eleven multiplies, few adds.
__device__ __forceinline__ uint64_t redc_words(uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3,
                                              uint64_t m, uint64_t mp) {
  const uint32_t m0 = (uint32_t)m, m1 = (uint32_t)(m >> 32);
  const uint32_t mp0 = (uint32_t)mp, mp1 = (uint32_t)(mp >> 32);
  const uint64_t qa = (uint64_t)t0 * mp0;
  const uint32_t q0 = (uint32_t)qa;
  const uint32_t q1 = (uint32_t)(qa >> 32) + t0 * mp1 + t1 * mp0;
  const uint64_t e = (uint64_t)q0 * m0 + t0;                  // low word 0
  const uint64_t f = (uint64_t)q1 * m0 + ((e >> 32) + t1);
  const uint64_t g = (uint64_t)q0 * m1 + (uint32_t)f;         // m1 < 2^23; low word 0
  const uint64_t hi = (uint64_t)q1 * m1 + (f >> 32) + (g >> 32);
  return (((uint64_t)t3 << 32) | t2) + hi;
}

__device__ __forceinline__ uint64_t isq55(uint64_t h, uint64_t m, uint64_t mp) {
  const uint32_t h0 = (uint32_t)h, h1 = (uint32_t)(h >> 32);  // h1 < 2^24
  const uint64_t a = (uint64_t)h0 * h0;
  const uint64_t x = (uint64_t)h0 * (h1 + h1) + (a >> 32);
  const uint64_t y = (uint64_t)h1 * h1 + (x >> 32);
  return redc_words((uint32_t)a, (uint32_t)x, (uint32_t)y, (uint32_t)(y >> 32), m, mp);
}

__device__ __forceinline__ uint64_t isq55_shift(uint64_t h, int c, uint64_t m, uint64_t mp) {
  const uint64_t gw = h << c;  // < 2^63
  const uint32_t h0 = (uint32_t)h, h1 = (uint32_t)(h >> 32);
  const uint32_t g0 = (uint32_t)gw, g1 = (uint32_t)(gw >> 32);
  const uint64_t a = (uint64_t)h0 * g0;
  const uint64_t x = (uint64_t)h0 * g1 + (a >> 32);
  const uint64_t x2 = (uint64_t)h1 * g0 + (uint32_t)x;
  const uint64_t y = (uint64_t)h1 * g1 + (x >> 32) + (x2 >> 32);
  return redc_words((uint32_t)a, (uint32_t)x2, (uint32_t)y, (uint32_t)(y >> 32), m, mp);
}

// Montgomery form of 2^E (= 2^E * 2^64 mod m) in [0, 2m) for m < 2^55: W-bit windows over the
// low bl bits of E (bl uniform per wave, E < 2^bl), starting from 2^64 mod m, the Montgomery form
// of 1; leading zero windows leave it unchanged.
template <int W>
__device__ __forceinline__ uint64_t ipow2_mont(uint64_t E, int bl, uint64_t m, uint64_t mp,
                                               uint64_t one) {
  int pos = ((bl + W - 1) / W) * W - W;  // start of the top (possibly partial) window
  uint64_t h = isq55_shift(one, (int)(E >> pos), m, mp);
  for (pos -= W; pos >= 0; pos -= W) {
#pragma unroll
    for (int i = 0; i < W - 1; ++i) h = isq55(h, m, mp);
    h = isq55_shift(h, (int)((E >> pos) & ((1u << W) - 1)), m, mp);
  }
  return h;
}

// Same for m < 2^63 with fully reduced residues and one bit per square.
__device__ __forceinline__ uint64_t ipow2_mont_wide(uint64_t E, uint64_t m, uint64_t mp,
                                                    uint64_t one) {
  uint64_t h = one;
  for (int pos = 63 - __builtin_clzll(E | 1); pos >= 0; --pos) {
    h = isq_shift(h, (int)((E >> pos) & 1), m, mp);
    if (h >= m) h -= m;
  }
  return h;
}

// Quotient words of floor(r 2^192 / m) from h = 2^(e+192) * 2^64 mod m in [0, 2m):
// s3 = REDC(h) = r 2^192 mod m, then REDC(s_i) = s_{i-1} with quotient word s_i * mp.
__device__ __forceinline__ void iquot192(uint64_t h, uint64_t m, uint64_t mp, uint64_t& q0,
                                         uint64_t& q1, uint64_t& q2) {
  if (h >= m) h -= m;
  const uint64_t s3 = mulhi64(h * mp, m) + (h != 0);
  q0 = s3 * mp;
  const uint64_t s2 = mulhi64(q0, m) + (s3 != 0);
  q1 = s2 * mp;
  const uint64_t s1 = mulhi64(q1, m) + (s2 != 0);
  q2 = s1 * mp;
}

}  // namespace clb

// clb_replay.hip -- reference replay: the rounding of the reference's kernels of all three
// precisions reproduced to the bit, for callers that need the reference's exact output bits rather
// than the exact sum the other kernels compute.
//
// The reference's 192-bit kernels keep, per work-item, two double-double sums over the
// work-item's strided terms (of r/m and of (r 2^96 mod m)/m, r = 1024^(d-k) mod m) and convert
// them to 192 bits once (Bellard.cl:771-823, :992-1084, :506-578); its tail kernel is
// double-double throughout (:1153-1253).  Everything after that is exact integer addition, so a
// replay must reproduce each work-item's double-double arithmetic operation by operation, and may
// add the work-items' 192-bit values in any order.  Here one lane is one work-item: a tile is 64
// consecutive work-items of one reference launch, tiles come from a queue in chunks, and each
// wave adds its lanes' values into the job's accumulator slot as the exact kernels do.  The
// 106-bit mode (precision 0) keeps double-doubles to the end, so its work-group trees and its
// 65536 group accumulators, updated launch by launch, are replayed in the reference's order too
// (k_r106_groups, k_r106_finish, k_r106_sum7).
//
// Contraction is off: each fused multiply-add in the double-double code below is one that the
// reference's OpenCL build makes (OpenCL C contracts a*b + c within one expression) and is written
// out; every other operation is rounded on its own, to nearest, as there.  The file is built in
// the device FP mode of the other kernels (IEEE mode bit clear, FP64 denormals flushed, NaNs
// assumed absent), which changes no result here: the mode bit only affects NaN handling and
// output modifiers, and every value is finite and either 0 or of magnitude above 2^-600, far
// from the denormal range.  The residues r are exact integers in the reference (its Montgomery
// exponentiations, Bellard.cl:395-461), so they come from this library's exact engines: the fp64
// engine of the exact kernels (doubling squares included) for m < 2^48, the integer engine above.
#include "clb_device.hpp"
#include "clb_internal.hpp"
#include "clb_wave.hpp"

namespace clb {
namespace {

struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {  // twosum, Bellard.cl:1-7
  const double x = a + b;
  const double t = x - a;
  return {x, (a - (x - t)) + (b - t)};
}

// twoproduct, :17-27: Dekker's product of the Veltkamp halves, its error term one expression whose
// four products the reference's build fuses, in the order (ah bh - p) + al bh + ah bl + al bl.
// Dekker's product is error-free (Dekker 1971; with no underflow or overflow every partial sum
// is exact, and the fused order above is his algorithm applied to (b, a)), so the error term is
// exactly a b - p, which one fma returns: the same two doubles.  The operands here are quotients
// of magnitude at least 2^-300 (or 0) and integers below 2^53, far from underflow and overflow.
__device__ __forceinline__ DD two_prod(double a, double b) {
  const double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {  // addcab, :29-35
  const DD c = two_sum(a.hi, b.hi);
  return two_sum(c.hi, (c.lo + a.lo) + b.lo);
}

__device__ __forceinline__ DD dd_neg(DD a) { return {-a.hi, -a.lo}; }

__device__ __forceinline__ DD div_nn(double x, double y) {  // nn_div, :99-106
  const double z1 = x / y;
  const DD c = two_prod(-z1, y);
  return two_sum(z1, ((c.hi + x) + c.lo) / y);
}

// x / y as the compiler lowers it on gfx950: div_scale of both operands, rcp and two Newton
// steps (recip), q = x r, its residual, div_fmas, div_fixup.  For y an integer in [1, 2^53) and
// x = +0 or |x| >= 2^-900, both div_scale steps return their input with VCC clear, div_fmas is
// then a plain fma and div_fixup returns its input (sign(x) ^ sign(y) is the sign of q, and a
// +0 numerator gives +0 either way), so with r = recip(y) the remaining three instructions give
// the same bits.  recip(y) is the exponent engine's reciprocal too, so one serves both.
__device__ __forceinline__ double div_r(double x, double y, double r) {
  const double q = x * r;
  return __builtin_fma(__builtin_fma(-y, q, x), r, q);
}

// fast_two_sum(a, b) = two_sum(a, b) when |b| <= |a|: both return RN(a + b) and the exact error
// (Dekker's Fast2Sum is error-free under that condition), in three instructions instead of six.
__device__ __forceinline__ DD fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}

// nn_div for a numerator in [0, y] and an integer modulus y in [1, 2^53), r = recip(y): its
// second numerator is +0 or an exact residual of magnitude >= ulp(x / y) >= 2^-106.  z1 is the
// correctly rounded x / y, so that residual is at most y ulp(z1) / 2 and the second quotient at
// most ulp(z1) <= |z1| (both are 0 when x is): the final two_sum is a fast_two_sum.
__device__ __forceinline__ DD div_nn_m(double x, double y, double r) {
  const double z1 = div_r(x, y, r);
  const DD c = two_prod(-z1, y);
  return fast_two_sum(z1, div_r((c.hi + x) + c.lo, y, r));
}

// nd_div, :89-96, and dd_div, :70-77: "- z1 * y.y" is fused with the sum before it.
__device__ __forceinline__ DD div_nd(double x, DD y) {
  const double z1 = x / y.hi;
  const DD c = two_prod(-z1, y.hi);
  return two_sum(z1, (__builtin_fma(-z1, y.lo, c.hi + x) + c.lo) / y.hi);
}

__device__ __forceinline__ DD div_dd(DD x, DD y) {
  const double z1 = x.hi / y.hi;
  const DD c = two_prod(-z1, y.hi);
  return two_sum(z1, ((__builtin_fma(-z1, y.lo, c.hi + x.hi) + x.lo) + c.lo) / y.hi);
}

__device__ __forceinline__ void fix_unit(DD& s) {  // sumfix2, :472-489
  double lo = 0.0;
  if (s.hi >= 1.0) lo = -1.0;
  if (s.hi <= -1.0) lo = 1.0;
  const DD t = two_sum(s.hi, lo);
  s = two_sum(t.hi, t.lo + s.lo);
}

// dd2_to_ulong3, :506-578: 52 + 12 bits of s0, 40 more, then 96 bits of s1 (48 + 48) with the
// two overlapping bits (30-31 of the middle word, 46-47 of s1's first 48) reconciled.
__device__ U192 to192(DD s0, DD s1) {
  const DD one = {1.0, 0.0};
  if (s0.hi < 0.0) s0 = dd_add(s0, one);
  if (s1.hi < 0.0) s1 = dd_add(s1, one);
  s0.hi *= 0x1p52;
  s0.lo *= 0x1p52;
  double t = __builtin_trunc(s0.hi);
  uint64_t u0 = (uint64_t)t * 4096u;
  s0 = two_sum(s0.hi - t, s0.lo);
  s0.hi *= 0x1p52;
  if (s0.hi < 0.0) {
    s0.hi += 0x1p52;
    u0 -= 4096u;
  }
  uint64_t ut = (uint64_t)s0.hi;
  u0 |= ut >> 40;
  uint64_t u1 = (ut & ((1ull << 40) - 1)) << 24;
  s1.hi *= 0x1p48;
  s1.lo *= 0x1p48;
  t = __builtin_trunc(s1.hi);
  ut = (uint64_t)t;
  s1 = two_sum(s1.hi - t, s1.lo);
  s1.hi *= 0x1p48;
  if (s1.hi < 0.0) {
    s1.hi += 0x1p48;
    ut -= 1;
  }
  const uint64_t u2 = ((ut & 0xffffull) << 48) + (uint64_t)s1.hi;
  const uint32_t chk0 = (uint32_t)((u1 >> 30) & 3u);
  const uint32_t chk1 = (uint32_t)(ut >> 46);
  if (chk0 != chk1) {
    const uint64_t last = u1;
    if (((chk0 + 1u) & 3u) == chk1) {
      u1 += 1ull << 30;
      if (u1 < last) u0++;
    }
    if (((chk0 - 1u) % 4u) == chk1) {
      u1 -= 1ull << 30;
      if (u1 > last) u0--;
    }
  }
  u1 = (u1 & 0xFFFFFFFF00000000ull) | (ut >> 16);
  return {u2, u1, u0};
}

// One term's update of a sum (:786-796, :1037-1048) from the residue x of 2^e mod m, |x| < m
// (r = x, or x + m when x < 0): while the sum is positive the reference divides m - r and negates
// the quotient, which keeps the sum in (-1, 1).  nn_div and addcab are odd-symmetric under round
// to nearest (every operation is), so dividing r - m instead gives the same sum: equal doubles,
// a zero at most differing in sign, which no later operation here turns into a nonzero
// difference.  The numerator is then one exact fma, x + c m with c = [x < 0] - [sum > 0], an
// integer below 2^53 in magnitude.  m is an exact integer held in a double, minv = recip(m).
__device__ __forceinline__ void dd_step(DD& s, double x, double m, double minv) {
  const double c = (x < 0.0 ? 1.0 : 0.0) - (s.hi > 0.0 ? 1.0 : 0.0);
  s = dd_add(s, div_nn_m(__builtin_fma(c, m, x), m, minv));
}

__device__ __forceinline__ DD dd_of(uint64_t v) {  // (double)v and (double)((long)v - (long)x)
  const double x = (double)v;
  return {x, (double)((int64_t)v - (int64_t)x)};
}

// The same update from an integer residue r in [0, m), m < 2^63.  Moduli of 2^53 and above occur
// only when 10 d + 9 >= 2^53, where the reference host builds its kernels with -D OVERDOUBLE
// (Bellard_192.hsp:472-474): the steered numerator and m then go into dd_div split into
// double-doubles (Bellard.cl:938-939, :957-962, :1040-1046).  Below 2^53 both parts' low halves
// are 0, and dd_div of such operands returns nn_div's doubles (its extra terms, z1 0 and + 0,
// change no value), so the build's choice only matters from 2^53 on.
__device__ __forceinline__ void dd_step_u(DD& s, uint64_t r, uint64_t m, double mf, double minv) {
  if (m >= (1ull << 53)) {
    const bool pos = s.hi > 0.0;
    DD q = div_dd(dd_of(pos ? m - r : r), dd_of(m));
    if (pos) q = dd_neg(q);
    s = dd_add(s, q);
  } else {
    dd_step(s, (double)r, mf, minv);
  }
}

// Window width of the fp64 engine for a kernel kind (0 for the integer engine).
template <int KIND>
constexpr int kWindow = KIND == kSplitW3 ? 3 : KIND == kSplitW2 ? 2 : KIND == kSplitW1 ? 1 : 0;

// Uniform start shift of the fp64 engine for one round of a wave: the smallest multiple s of W
// with (E0 >> s) <= tmax = 49 + floor(log2 m0), E0 the wave's largest exponent and m0 its
// smallest modulus (both lane 0's).  Every lane's start 2^t, t = e >> s <= tmax, then has
// 2^t / m <= 2^49, which fpow2's one-step reduction of 2^t needs (|x| <= 0.69 m).
template <int W>
__device__ __forceinline__ int start_shift(uint64_t E0, uint64_t m0) {
  const int tmax = 49 + 63 - __builtin_clzll(m0);
  const int bl = 64 - __builtin_clzll(E0 | 1);
  const int bt = 32 - __builtin_clz((uint32_t)tmax);
  int s = bl > bt ? W * ((bl - bt + W - 1) / W) : 0;  // now E0 >> s < 2^bt
  if ((E0 >> s) > (uint64_t)tmax) s += W;             // E0 >> s < 2^(bt - 1) <= tmax
  return s;
}

// 2^e mod m, balanced (|x| < m), for lane `lane` of a wave whose lane 0 has the exponent E0 and
// the smallest modulus m0, e = E0 - 10 lane (consecutive k).  For E0 >= 1664 this is the exact
// kernels' form (fpow2_split): a uniform H = (E0 - 640) >> 10 whose bits take plain or doubling
// squares, then the lane's v = ((E0 - 640) & 1023) + 640 - 10 lane in [10, 1663] in W-bit windows;
// the result is within 0.92 m (0.67 m after the final plain square of W = 3 and 1).  Smaller
// exponents take per-lane windows from the start shift above (within 0.61 m).
template <int W>
__device__ __forceinline__ double pow2_lane(uint64_t E0, uint32_t lane, uint64_t m0, double m,
                                            double minv) {
  if (E0 >= 1664) {
    const uint64_t F = E0 - 640;
    const int tmax = 49 + 63 - __builtin_clzll(m0);
    return fpow2_split<W>(F >> 10, (uint32_t)(F & 1023) + 640u - 10u * lane, tmax, m, minv);
  }
  return fpow2<W>(E0 - 10ull * lane, start_shift<W>(E0, m0), m, minv);
}

template <int KIND>
constexpr bool kIntEngine = KIND == kPlainW3 || KIND == kPlainWide;

// The int engine's residues (kPlainW3: odd m < 2^55, 3-bit windows, lazy residues; kPlainWide:
// m < 2^63, one bit per square, reduced), in [0, m): x = 2^e mod m and, with b, x 2^96 mod m
// (:798-799, :1052-1053), e = E0 - 10 lane.  From the Montgomery form h = 2^(e + 64) mod m:
// x = REDC(h), b = REDC(h (2^96 mod m)), 2^96 mod m being the Montgomery form of 2^32.
template <int KIND>
__device__ __forceinline__ void int_residues(uint64_t E0, uint32_t lane, uint64_t m, uint64_t& x,
                                             uint64_t* b) {
  const uint64_t mp = neg_inv64(m);
  const uint64_t one = (0ull - m) % m;
  uint64_t h, c = 0;
  if constexpr (KIND == kPlainW3) {
    const int bl = 64 - __builtin_clzll(E0);  // every lane's e < 2^bl
    h = ipow2_mont<3>(E0 - 10ull * lane, bl, m, mp, one);
    if (b) c = ipow2_mont<3>(32, 6, m, mp, one);
  } else {
    h = ipow2_mont_wide(E0 - 10ull * lane, m, mp, one);
    if (b) c = ipow2_mont_wide(32, m, mp, one);
  }
  x = redc(0, h, m, mp);
  if (x >= m) x -= m;
  if (b) {
    uint64_t v = redc(mulhi64(h, c), h * c, m, mp);
    if (v >= m) v -= m;
    *b = v;
  }
}

// The fp64 engine's residues (KIND kSplitW3/W2/W1, m < 2^42 / 2^46 / 2^48; E0 and m0 = lane 0's
// exponent and modulus, uniform; mf = m, minv = recip(mf)): x = 2^e mod m balanced by pow2_lane
// (|x| < 0.92 m), and b = x 2^96 mod m by two exact shift-and-reduce steps of 48 bits
// (|x 2^48 / m - q| <= 0.5 + 2^48 1.5 2^-52 < 0.6, so |b| < 0.6 m and every intermediate is an
// exact integer below 2^53).
template <int KIND>
__device__ __forceinline__ void fp_residues(uint64_t E0, uint32_t lane, uint64_t m0, double mf,
                                            double minv, double& x, double& b) {
  x = pow2_lane<kWindow<KIND>>(E0, lane, m0, mf, minv);
  b = x;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double z = __builtin_amdgcn_ldexp(b, 48);
    b = __builtin_fma(-__builtin_rint(z * minv), mf, z);
  }
}

// Work-item idx = 64 tile + lane of a replayed launch (:783-810, :1012-1078); idx < the launch's
// work-items.  Its terms are k = offset + idx + i gsize < k_max; lane 0 of the tile has the
// largest exponent of each round, which fixes the uniform window count.
template <int KIND>
__device__ U192 replay_item(const ReplayJobDev& jb, uint64_t tile, int lane) {
  DD s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
  const uint64_t d = jb.d, k_max = jb.k_max, gsize = jb.gsize, den0 = jb.den0, den1 = jb.den1;
  const uint64_t k_lane0 = jb.offset + tile * kWave;
  // The lane's modulus advances by den0 gsize per term: for the fp64 engine one exact add of
  // doubles (every m used is below 2^48), for the int engine a 64-bit add and a conversion.
  uint64_t m = den0 * (k_lane0 + (uint64_t)lane) + den1;
  double mf = (double)m;
  const uint64_t mstep = den0 * gsize;
  const double mstepf = (double)mstep;
  for (uint64_t kw = k_lane0; kw < k_max; kw += gsize) {
    const uint64_t k = kw + (uint64_t)lane;
    if (k < k_max) {
      if constexpr (kIntEngine<KIND>) mf = (double)m;
      const double minv = recip(mf);  // the exponent engine's and the divisions'
      if constexpr (kIntEngine<KIND>) {
        uint64_t x, b;
        int_residues<KIND>(10ull * (d - kw), (uint32_t)lane, m, x, &b);
        dd_step_u(s0, x, m, mf, minv);
        dd_step_u(s1, b, m, mf, minv);
      } else {
        double x, b;
        fp_residues<KIND>(10ull * (d - kw), (uint32_t)lane, den0 * kw + den1, mf, minv, x, b);
        dd_step(s0, x, mf, minv);
        dd_step(s1, b, mf, minv);
      }
    }
    if constexpr (kIntEngine<KIND>)
      m += mstep;
    else
      mf += mstepf;
  }
  // The reference's k after its loop has the parity of offset + idx (gsize is even).
  if ((jb.numesign + (k_lane0 + (uint64_t)lane) % 2) % 2 == 1) {
    s0 = dd_neg(s0);
    s1 = dd_neg(s1);
  }
  return to192(s0, s1);
}

__device__ uint64_t pow2_mod_small(int e, uint64_t m) {  // 2^e mod m, m < 2^63, by doubling
  uint64_t x = 1 % m;
  for (int i = 0; i < e; ++i) {
    x <<= 1;
    if (x >= m) x -= m;
  }
  return x;
}

// Sglobal_after_192, :1153-1253: k = d + i, i = 19..0, into s0 with numerators 2^(-10 i);
// i = 19..10 into s1 with 2^(96 - 10 i); i = 9..0 into s1 with the residues 2^(96 - 10 i) mod m
// (exact in the reference) and sumfix2 after each.
__device__ U192 replay_tail(const TailDev& t) {
  DD s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
  double n0 = 1.0, n1 = 0x1p96;
  for (int i = 0; i < 19; ++i) {
    n0 /= 1024.0;
    n1 /= 1024.0;
  }
  int64_t i = 19;
  for (; i >= 0; --i) {
    const uint64_t k = t.d + (uint64_t)i;
    DD q = div_nd(n0, dd_of(t.den0 * k + t.den1));
    if (k % 2 == 1) q = dd_neg(q);
    s0 = dd_add(s0, q);
    n0 *= 1024.0;
  }
  for (i = 19; i >= 10; --i) {
    const uint64_t k = t.d + (uint64_t)i;
    DD q = div_nd(n1, dd_of(t.den0 * k + t.den1));
    if (k % 2 == 1) q = dd_neg(q);
    s1 = dd_add(s1, q);
    n1 *= 1024.0;
  }
  for (; i >= 0; --i) {
    const uint64_t k = t.d + (uint64_t)i;
    const uint64_t m = t.den0 * k + t.den1;
    const uint64_t nm = pow2_mod_small(96 - 10 * (int)i, m);
    DD q = m > 4294967296ull ? div_dd(dd_of(nm), dd_of(m)) : div_nn((double)nm, (double)m);
    if (k % 2 == 1) q = dd_neg(q);
    s1 = dd_add(s1, q);
    fix_unit(s1);
  }
  if (t.numesign == 1) {
    s0 = dd_neg(s0);
    s1 = dd_neg(s1);
  }
  return to192(s0, s1);
}

__device__ __forceinline__ void flush(U192& v, int lane, unsigned long long* slot) {
  wave_sum192(v);
  if (lane == 0) atomic_add192(slot, v);
  v = {0, 0, 0};
}

// ------------------------------------------------------------------ 106-bit mode

__device__ __forceinline__ DD dd_sub(DD a, DD b) {  // subcab, :37-43
  const DD c = two_sum(a.hi, -b.hi);
  return two_sum(c.hi, (c.lo + a.lo) - b.lo);
}

__device__ __forceinline__ void fix_one(DD& s) {  // sumfix, :491-501
  if (s.hi >= 1.0) {
    const DD t = two_sum(s.hi, -1.0);
    s = two_sum(t.hi, t.lo + s.lo);
  }
}

// The 256 work-items of one work-group reduced in LDS the reference's way (shared_reduction_dd,
// :583-597, and Ssum's tree, :1274-1297): p[l] = sumfix2(p[l] + p[l + i]), i = 128 .. 1.
__device__ __forceinline__ void tree256(DD* p, uint32_t tid) {
  for (uint32_t i = 128; i > 0; i >>= 1) {
    __syncthreads();
    if (tid < i) {
      DD t = dd_add(p[tid], p[tid + i]);
      fix_unit(t);
      p[tid] = t;
    }
  }
  __syncthreads();
}

// Work-item 256 g + tid of one 106-bit-mode launch: Sglobal32mtg (:729-752, sumfix after each
// term) or Sglobal64mtg (:914-975, the steered numerator), and the sign of its terms.
template <int KIND>
__device__ DD item106(const R106JobDev& jb, uint64_t offset, uint64_t k_max, uint64_t gsize,
                      uint32_t g, uint32_t tid, bool is64) {
  DD s = {0.0, 0.0};
  const uint64_t idx = (uint64_t)g * 256 + tid;
  const uint64_t lane = tid & (kWave - 1);
  // The wave's first work-item, made scalar (readfirstlane): the compiler cannot tell that
  // tid - lane is uniform, and per-wave values (exponent length, start shift) would take VALU.
  const uint64_t wave0 =
      (uint64_t)g * 256 + __builtin_amdgcn_readfirstlane(tid & ~(uint32_t)(kWave - 1));
  // m advances by den0 gsize per term (as in replay_item).
  uint64_t m = jb.den0 * (offset + idx) + jb.den1;
  double mf = (double)m;
  const uint64_t mstep = jb.den0 * gsize;
  const double mstepf = (double)mstep;
  for (uint64_t kw = offset + wave0; kw < k_max; kw += gsize) {  // kw: the wave's lane 0
    const uint64_t k = kw + lane;
    if (k < k_max) {
      if constexpr (kIntEngine<KIND>) mf = (double)m;
      const double minv = recip(mf);
      if constexpr (kIntEngine<KIND>) {
        uint64_t r;
        int_residues<KIND>(10ull * (jb.d - kw), (uint32_t)lane, m, r, nullptr);
        if (is64) {
          dd_step_u(s, r, m, mf, minv);
        } else {  // m < 2^32
          s = dd_add(s, div_nn_m((double)r, mf, minv));
          fix_one(s);
        }
      } else {
        // x = 2^e mod m alone, balanced (the fp64 engine)
        const double x = pow2_lane<kWindow<KIND>>(10ull * (jb.d - kw), (uint32_t)lane,
                                                   jb.den0 * kw + jb.den1, mf, minv);
        if (is64) {
          dd_step(s, x, mf, minv);
        } else {
          s = dd_add(s, div_nn_m(x < 0.0 ? x + mf : x, mf, minv));
          fix_one(s);
        }
      }
    }
    if constexpr (kIntEngine<KIND>)
      m += mstep;
    else
      mf += mstepf;
  }
  if ((jb.numesign + (offset + idx) % 2) % 2 == 1) s = dd_neg(s);
  return s;
}

// Sglobal_after (:1118-1150): k = d .. d+9, numerators 2^(-10 i), the formula's sign.
__device__ DD tail106(const R106JobDev& t) {
  DD s = {0.0, 0.0};
  double n = 1.0;
  for (uint64_t i = 0; i < 10; ++i) {
    const uint64_t k = t.d + i;
    DD q = div_nd(n, dd_of(t.den0 * k + t.den1));
    if (k % 2 == 1) q = dd_neg(q);
    s = dd_add(s, q);
    n /= 1024.0;
  }
  if ((t.numesign + 2) % 2 == 1) s = dd_neg(s);
  return s;
}

}  // namespace

// Group g (blockIdx.x) of job blockIdx.y: the job's launches in the reference's order (GWS * 16
// terms below loopmax32, GWS * 32 above, KernelFuncs.as:31-53), each reduced over the group's 256
// work-items and added into the group's accumulator (:760-764, :979-983).  A group beyond every
// launch's work-items keeps its accumulator: adding the zero tree of its idle work-items, then
// sumfix2, leaves a double-double already in (-1, 1) unchanged, and only bigSum[0] (after the
// tail) can lie outside, whose group always has work-items.
template <int KIND>
__global__ void __launch_bounds__(256) k_r106_groups(const R106JobDev* __restrict__ jobs,
                                                    uint64_t gsize) {
  __shared__ DD p[256];
  const R106JobDev jb = jobs[block_y()];
  const uint32_t g = block_x(), tid = thread_x();
  const uint64_t span = jb.k1 > jb.k0 ? jb.k1 - jb.k0 : 0;
  const uint64_t first_items = span < gsize ? span : gsize;
  if ((uint64_t)g * 256 >= first_items && !(jb.single && g == 0)) return;
  double* acc_ptr = reinterpret_cast<double*>(jb.bigsum) + 2 * (size_t)g;
  DD acc = {acc_ptr[0], acc_ptr[1]};
  uint64_t offset = jb.k0, k_max = jb.k0;
  const uint64_t x32 = ((1ull << 32) - jb.den1 - 1) / jb.den0 + 1;
  const uint64_t lm32 = x32 < jb.k1 ? x32 : jb.k1;
  if (jb.single) {
    // One launch over [k0, k1) of the kernel for m(k0) (clb_sglobal).  Group 0 runs even when the
    // range is empty: sumfix2 of its sum then brings a tail-raised bigSum[0] back into (-1, 1),
    // as the reference's launch of idle work-items does.
    p[tid] = item106<KIND>(jb, jb.k0, jb.k1, gsize, g, tid, jb.k0 >= x32);
    tree256(p, tid);
    if (tid == 0) {
      DD t = dd_add(acc, p[0]);
      fix_unit(t);
      acc = t;
    }
    if (tid == 0) {
      acc_ptr[0] = acc.hi;
      acc_ptr[1] = acc.lo;
    }
    return;
  }
  for (int phase = 0; phase < 2; ++phase) {
    const uint64_t lim = phase == 0 ? lm32 : jb.k1;
    const uint64_t step = gsize * (phase == 0 ? 16 : 32);
    while (offset < lim) {
      k_max = k_max + step < lim ? k_max + step : lim;
      const uint64_t items = k_max - offset < gsize ? k_max - offset : gsize;
      if ((uint64_t)g * 256 < items) {
        p[tid] = item106<KIND>(jb, offset, k_max, gsize, g, tid, phase == 1);
        tree256(p, tid);
        if (tid == 0) {
          DD t = dd_add(acc, p[0]);
          fix_unit(t);
          acc = t;
        }
        __syncthreads();
      }
      offset = k_max;
    }
  }
  if (tid == 0) {
    acc_ptr[0] = acc.hi;
    acc_ptr[1] = acc.lo;
  }
}

// Job blockIdx.x: the tail into bigSum[0] without sumfix2 (:1149), then (with ssum) Ssum
// 65536 -> 256 -> 1 (KernelFuncs.as:62-65); bigSum[0] and the job's part receive the total.
__global__ void __launch_bounds__(256) k_r106_finish(const R106JobDev* __restrict__ jobs) {
  __shared__ DD p[256];
  __shared__ DD small[256];
  const R106JobDev jb = jobs[block_x()];
  const uint32_t tid = thread_x();
  double* bs = reinterpret_cast<double*>(jb.bigsum);
  if (tid == 0 && jb.tail) {
    const DD b0 = dd_add({bs[0], bs[1]}, tail106(jb));
    bs[0] = b0.hi;
    bs[1] = b0.lo;
  }
  double* part = reinterpret_cast<double*>(jb.part);
  if (!jb.ssum) {  // uniform per block: bigSum[0] as it stands
    if (tid == 0) {
      part[0] = bs[0];
      part[1] = bs[1];
    }
    return;
  }
  __syncthreads();
  for (uint32_t b = 0; b < 256; ++b) {
    const size_t i = 2 * ((size_t)b * 256 + tid);
    p[tid] = {bs[i], bs[i + 1]};
    tree256(p, tid);
    if (tid == 0) small[b] = p[0];
    __syncthreads();
  }
  p[tid] = small[tid];
  tree256(p, tid);
  if (tid == 0) {
    bs[0] = part[0] = p[0].hi;
    bs[1] = part[1] = p[0].lo;
  }
}

// AddAns (Sum7, :1324-1356) per result group g: for each of its slots in order, part <- frac(numer
// part) (written to slot_val) and ans <- sumfix(ans + part).
__global__ void k_r106_sum7(const double* parts, const uint32_t* first, const uint32_t* count,
                            const uint8_t* shift, double* ans, double* slot_val,
                            uint32_t ngroups) {
  const uint32_t g = block_x() * block_dim_x() + thread_x();
  if (g >= ngroups) return;
  DD a = {ans[2 * g], ans[2 * g + 1]};
  for (uint32_t i = 0; i < count[g]; ++i) {
    const size_t s = first[g] + i;
    DD v = {parts[2 * s], parts[2 * s + 1]};
    if (v.hi < 0.0) v = dd_add(v, {1.0, 0.0});
    const double numer = (double)(1u << shift[s]);
    v.hi *= numer;
    v.lo *= numer;
    v = dd_sub(v, {__builtin_trunc(v.hi), 0.0});
    slot_val[2 * s] = v.hi;
    slot_val[2 * s + 1] = v.lo;
    DD t = dd_add(a, v);
    fix_one(t);
    a = t;
  }
  ans[2 * g] = a.hi;
  ans[2 * g + 1] = a.lo;
}

// Persistent replay kernel: tails first (wave w: tails w, w + nwaves, ..., one lane each), then
// chunks of consecutive replay tiles from the queue.  A wave's tiles only move forward, so its
// job index does too; it adds its lanes' values into the job's slot when the job changes.
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_replay(LaunchDev* __restrict__ hdr) {
  unsigned long long* __restrict__ acc = hdr_acc(hdr);
  const ReplayJobDev* __restrict__ jobs = hdr_rjobs(hdr);
  const int njobs = (int)hdr->nrjobs;
  const uint64_t total = hdr->rtotal_tiles;
  const uint64_t chunk = hdr->rchunk;
  const uint64_t nwaves = (uint64_t)grid_dim_x() * (kBlock / kWave);
  const uint64_t wave_id = (uint64_t)block_x() * (kBlock / kWave) +
                           (uint64_t)__builtin_amdgcn_readfirstlane(thread_x() / kWave);
  const int lane = (int)(thread_x() & (kWave - 1));
  const uint32_t replicas = hdr->replicas;
  const uint32_t rep = (uint32_t)wave_id & (replicas - 1);
  const TailDev* __restrict__ tails = hdr_rtails(hdr);
  for (uint64_t t = wave_id; t < hdr->nrtails; t += nwaves) {
    if (lane == 0)
      atomic_add192(acc + ((size_t)tails[t].slot * replicas + rep) * kAccWords,
                    replay_tail(tails[t]));
  }
  int cur = -1;
  ReplayJobDev jb{};
  U192 sum = {0, 0, 0};
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&hdr->rcounter, (unsigned long long)chunk);
    const uint64_t f0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)base);
    if (f0 >= total) break;
    const uint64_t f1 = f0 + chunk < total ? f0 + chunk : total;
    for (uint64_t f = f0; f < f1; ++f) {
      int lo = cur < 0 ? 0 : cur, hi = njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].flat_begin <= f)
          lo = mid;
        else
          hi = mid - 1;
      }
      if (lo != cur) {
        if (cur >= 0) flush(sum, lane, acc + ((size_t)jb.slot * replicas + rep) * kAccWords);
        cur = lo;
        jb = jobs[lo];
      }
      const uint64_t tile = f - jb.flat_begin;
      const uint64_t span = jb.k_max - jb.offset;
      const uint64_t items = span < jb.gsize ? span : jb.gsize;
      if (tile * kWave + (uint64_t)lane < items) {
        const U192 v = replay_item<KIND>(jb, tile, lane);
        add192(sum, v.w0, v.w1, v.w2);
      }
    }
  }
  if (cur >= 0) flush(sum, lane, acc + ((size_t)jb.slot * replicas + rep) * kAccWords);
}

template <int KIND>
static int replay_blocks() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_replay<KIND>, kBlock, 0) != hipSuccess)
    n = 0;
  return n;
}

int replay_blocks_per_cu(int kind) {
  switch (kind) {
    case kSplitW3: return replay_blocks<kSplitW3>();
    case kSplitW2: return replay_blocks<kSplitW2>();
    case kSplitW1: return replay_blocks<kSplitW1>();
    case kPlainW3: return replay_blocks<kPlainW3>();
    case kPlainWide: return replay_blocks<kPlainWide>();
    default: return 0;
  }
}

hipError_t launch_replay106(int kind, const R106JobDev* jobs, uint32_t njobs, uint64_t gsize,
                            uint32_t ngroups_max, hipStream_t stream) {
  if (njobs == 0) return hipSuccess;
  if (ngroups_max > 0) {
    const dim3 g(ngroups_max, njobs), b(256);
    const uint64_t gs = gsize;
    switch (kind) {
      case kSplitW3: hipLaunchKernelGGL(k_r106_groups<kSplitW3>, g, b, 0, stream, jobs, gs); break;
      case kSplitW2: hipLaunchKernelGGL(k_r106_groups<kSplitW2>, g, b, 0, stream, jobs, gs); break;
      case kSplitW1: hipLaunchKernelGGL(k_r106_groups<kSplitW1>, g, b, 0, stream, jobs, gs); break;
      case kPlainW3: hipLaunchKernelGGL(k_r106_groups<kPlainW3>, g, b, 0, stream, jobs, gs); break;
      case kPlainWide:
        hipLaunchKernelGGL(k_r106_groups<kPlainWide>, g, b, 0, stream, jobs, gs);
        break;
      default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_r106_finish, dim3(njobs), dim3(256), 0, stream, jobs);
  return hipGetLastError();
}

hipError_t launch_replay106_sum7(const double* parts, const uint32_t* first, const uint32_t* count,
                                 const uint8_t* shift, double* ans, double* slot_val,
                                 uint32_t ngroups, hipStream_t stream) {
  if (ngroups == 0) return hipSuccess;
  hipLaunchKernelGGL(k_r106_sum7, dim3((ngroups + 63) / 64), dim3(64), 0, stream, parts, first,
                     count, shift, ans, slot_val, ngroups);
  return hipGetLastError();
}

// This file's own start-up check of the FP64 output modifier (its code object is built apart from
// clb_kernels.hip's): one lane runs fsq_bits on 3 mod 1000003 with the exponent bit set (18) and
// clear (9).
__global__ void k_replay_selfcheck(uint64_t* out) {
  if (thread_x() != 0) return;
  const double m = 1000003.0;
  const double minv = recip(m);
  out[0] = (uint64_t)(int64_t)fsq_bits(3.0, 2ull, 1u, m, minv);
  out[1] = (uint64_t)(int64_t)fsq_bits(3.0, 1ull, 1u, m, minv);
}

hipError_t launch_replay_selfcheck(uint64_t* out2, hipStream_t stream) {
  hipLaunchKernelGGL(k_replay_selfcheck, dim3(1), dim3(64), 0, stream, out2);
  return hipGetLastError();
}

hipError_t launch_replay(int kind, LaunchDev* hdr, int grid, hipStream_t stream) {
  const dim3 g(grid), b(kBlock);
  switch (kind) {
    case kSplitW3: hipLaunchKernelGGL(k_replay<kSplitW3>, g, b, 0, stream, hdr); break;
    case kSplitW2: hipLaunchKernelGGL(k_replay<kSplitW2>, g, b, 0, stream, hdr); break;
    case kSplitW1: hipLaunchKernelGGL(k_replay<kSplitW1>, g, b, 0, stream, hdr); break;
    case kPlainW3: hipLaunchKernelGGL(k_replay<kPlainW3>, g, b, 0, stream, hdr); break;
    case kPlainWide: hipLaunchKernelGGL(k_replay<kPlainWide>, g, b, 0, stream, hdr); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace clb

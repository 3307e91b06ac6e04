// clb_kernels.hip -- HIP kernels for gfx950: Bellard head terms, tail terms and the combine.
//
// Work decomposition.  A launch holds jobs (formula j, loop count d, head range [k0, k1)).  Each
// job's range is cut into tiles of 64 * TPL consecutive k aligned to multiples of the tile size;
// the tiles of all jobs form one flat index space that a persistent grid consumes, in static
// per-wave ranges for small launches and otherwise in chunks of consecutive tiles taken from a
// global counter (k_head).  Lane L of a tile evaluates k = base + L + 64 i,
// i < TPL, so every lane sees one parity of k and hence one sign (KernelFuncs.as:10 and the
// (numesign + k) % 2 rule of Bellard_div2by1.cl:335).  Each lane keeps the 192-bit sum Q of its
// terms; when the wave moves to another job (and at the end) the lane applies its sign (-Q - n for
// n negative terms, Bellard_div2by1.cl:335-341), the wave sums its 64 lanes with cross-lane
// shuffles, and lane 0 adds the six 32-bit limbs into the job's accumulator slot with 64-bit
// atomics.  Integer sums are exact, so the result does not depend on the schedule.
#include "clb_device.hpp"
#include "clb_internal.hpp"
#include "clb_wave.hpp"

namespace clb {

// Sign, wave sum and atomic add of the lanes' job sums.
__device__ __forceinline__ void flush_job(U192 v, uint64_t nterms, uint32_t numesign, int lane,
                                          unsigned long long* slot) {
  if ((numesign + (uint32_t)lane) & 1u) neg192_minus(v, nterms);
  wave_sum192(v);
  if (lane == 0) atomic_add192(slot, v);
}

// ------------------------------------------------------------------ per-tile term evaluation

// fp64 engine.  Per lane per tile at most TPL <= 16 terms, so the four digit sums stay exact
// integers below 2^53 and are converted to the lane's 192-bit sum once per tile.
// Lane L's exponent is E - 10 L with E the exponent of lane 0 (uniform): for E >= 1664 it is
// evaluated as 2^(1024 H + v) with H = (E - 640) >> 10 uniform and v = ((E - 640) & 1023) + 640
// - 10 L in [10, 1663] per lane (fpow2_split); smaller exponents take per-lane windows.
// 1/m of the next term (m + 64 den0) comes from this term's reciprocal by Newton steps
// r <- r + r (1 - m r), each squaring the relative error: the step 64 den0 <= 640 leaves it below
// 640 / m, so one step lands within an ulp of 1/m once m >= 2^37 and two steps once m >= 2^23;
// smaller moduli take the hardware estimate and two steps (recip).  nrecip (0, 1 or 2) is the
// tile's uniform choice for terms 1..TPL-1, from the modulus of its second term.
// kFast: every term of the tile lies in the job and has E >= 1664, and m >= 2^37 from the second
// term on, or m >= 2^23 in 16-term tiles (all tiles but the first ~10^6 k of a job and the last
// ~170 k before d), so the term loop carries no range check, no exponent-form test, no per-term
// count and no choice of reciprocal update: kRecip is nrecip, fixed.
template <int W, int TPL, bool kFast, int kRecip>
__device__ __forceinline__ void split_terms(const JobDev& jb, uint64_t kbase, int lane, int nt,
                                            int nrecip, uint64_t E0, int tmax, double (&f)[4],
                                            uint32_t& corr, uint64_t& nterms) {
  if (kFast) {
    nrecip = kRecip;
    nt = TPL;  // a constant trip count keeps the loop test scalar
  }
  const uint32_t ten_lane = 10u * (uint32_t)lane;
  double m = (double)(jb.den0 * (kbase + (uint64_t)lane) + jb.den1);
  const double mstep = (double)(jb.den0 * (uint64_t)kWave);
  double minv = recip(m);
  for (int i = 0; i < nt; ++i) {
    const uint64_t k = kbase + (uint64_t)kWave * (uint64_t)i + (uint64_t)lane;
    if (kFast || (k >= jb.k0 && k < jb.k1)) {
      const uint64_t E = E0 - 10ull * kWave * (uint64_t)i;
      double rem;
      if (kFast || E >= 1664) {
        const uint64_t F = E - 640;
        const uint32_t vs = (uint32_t)(F & 1023) + 640u;  // uniform part of v
        rem = fpow2_split<W>(F >> 10, vs - ten_lane, tmax, m, minv);
      } else {
        const int bl = 64 - __builtin_clzll(E);
        const int s = bl > 3 ? W * ((bl - 3) / W) : 0;
        rem = fpow2<W>(E - ten_lane, s, m, minv);
      }
      f[0] += fdigit(rem, m, minv);
      f[1] += fdigit(rem, m, minv);
      f[2] += fdigit(rem, m, minv);
      f[3] += fdigit(rem, m, minv);
      corr += rem >= m ? 1u : 0u;          // lazy remainder in [m, 2m): the last digit is one more
      if (!kFast) nterms += (m > 1.0);     // m = 1 gives r = 0 and a zero term of either sign
    }
    m += mstep;
    if (nrecip == 0) {
      minv = recip(m);
    } else {
      double e = __builtin_fma(-m, minv, 1.0);
      minv = __builtin_fma(minv, e, minv);
      if (nrecip == 2) {
        e = __builtin_fma(-m, minv, 1.0);
        minv = __builtin_fma(minv, e, minv);
      }
    }
  }
}

template <int W, int TPL>
__device__ __forceinline__ void tile_split(const JobDev& jb, uint64_t kbase, int lane, U192& acc,
                                           uint64_t& nterms) {
  static_assert(TPL >= 1 && TPL <= kMaxTermsPerLane, "digit sums must stay below 2^53");
  // Per-tile uniform facts, so that the term loop compares only 32-bit scalars:
  const uint64_t left = (jb.k1 - kbase + kWave - 1) / kWave;  // terms of lane 0 before k1
  const int nt = left < (uint64_t)TPL ? (int)left : TPL;
  const bool full = kbase >= jb.k0 && kbase + tile_k(TPL) <= jb.k1;  // every lane and term valid
  const uint64_t m1 = jb.den0 * (kbase + kWave) + jb.den1;  // smallest m of terms i >= 1
  const int nrecip = m1 >= (1ull << 37) ? 1 : m1 >= (1ull << 23) ? 2 : 0;
  const uint64_t E0 = 10ull * (jb.d - kbase);                    // exponent of lane 0, term 0
  // every term of the tile in 1024 H + v form
  const bool big = E0 - 10ull * kWave * (uint64_t)(nt - 1) >= 1664;
  const int tmax = 49 + 63 - __builtin_clzll(jb.den0 * kbase + jb.den1);  // smallest m of the tile
  double f[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t corr = 0;
  // A second fast loop costs the short-tile kernels (4 and 1 terms per lane, fully unrolled, for
  // launches too small to fill the GPU with 16-term tiles, where the moduli are small anyway)
  // registers and spills: there the two-step tiles take the general loop.
  constexpr bool kTwoStepFast = TPL == kMaxTermsPerLane;
  if (full && big && (nrecip == 1 || (kTwoStepFast && nrecip == 2))) {
    if (nrecip == 1)
      split_terms<W, TPL, true, 1>(jb, kbase, lane, nt, nrecip, E0, tmax, f, corr, nterms);
    else if constexpr (kTwoStepFast)
      split_terms<W, TPL, true, 2>(jb, kbase, lane, nt, nrecip, E0, tmax, f, corr, nterms);
    nterms += TPL;  // every term counted; k = 0 with den1 = 1 is the only m = 1
    if (kbase == 0 && lane == 0 && jb.den1 == 1) nterms -= 1;
  } else {
    split_terms<W, TPL, false, 0>(jb, kbase, lane, nt, nrecip, E0, tmax, f, corr, nterms);
  }
  // Digit i has weight 2^(144 - 48 i).  f[0] may be negative (the residue is balanced, |x| up to
  // 1.34 m, so |f[0]| < 16 1.34 2^48 < 2^53): only its low 48 bits matter mod 2^192, and its two's
  // complement carries them.  Digits 1..3 are in [0, 2^49).
  const uint64_t u3 = f2u64(f[3]) + corr;
  const uint64_t u2 = f2u64(f[2]);
  const uint64_t u1 = f2u64(f[1]);
  const uint64_t u0 = f2i64(f[0]);
  add192(acc, u3, 0, 0);
  add192(acc, u2 << 48, u2 >> 16, 0);
  add192(acc, 0, u1 << 32, u1 >> 32);
  add192(acc, 0, 0, u0 << 16);
}

// int engine: Montgomery form of 2^(e+192), then three exact-division quotient words.  The
// Montgomery form of 1, 2^64 mod m, is carried along the lane's moduli (m += 64 den0 per term)
// once m >= 2^37: with 2^64 = qq m + one, 2^64 = qq (m + s) + (one - qq s) and qq s < m, so the
// next residue is one - qq s, or that plus m + s (and qq - 1) when it is negative; one 64-bit
// division per lane per tile instead of one per term.  Smaller moduli (the first ~10^10 k) divide
// per term.
template <int KIND, int TPL>
__device__ __forceinline__ void tile_plain(const JobDev& jb, uint64_t kbase, int lane, U192& acc,
                                           uint64_t& nterms) {
  constexpr bool kWide = KIND == kPlainWide;
  uint64_t k = kbase + (uint64_t)lane;
  uint64_t m = jb.den0 * k + jb.den1;
  uint64_t e = 10ull * (jb.d - k);
  const uint64_t mstep = jb.den0 * (uint64_t)kWave;
  const bool carry_one = jb.den0 * kbase + jb.den1 >= (1ull << 37);  // uniform per tile
  uint64_t one = 0;
  uint32_t qq = 0;  // floor(2^64 / m) < 2^27
  if (carry_one) {
    const uint64_t q64 = UINT64_MAX / m;  // m odd > 1, so 2^64 mod m != 0
    qq = (uint32_t)q64;
    one = UINT64_MAX - q64 * m + 1;
  }
  for (int i = 0; i < TPL; ++i) {
    const uint64_t kw = kbase + (uint64_t)kWave * (uint64_t)i;
    if (kw >= jb.k1) break;
    if (k >= jb.k0 && k < jb.k1) {
      const uint64_t mp = neg_inv64(m);
      const uint64_t mont1 = carry_one ? one : (0ull - m) % m;  // 2^64 mod m
      const uint64_t E = e + 192;
      uint64_t h;
      if constexpr (kWide) {
        h = ipow2_mont_wide(E, m, mp, mont1);
      } else {
        const int bl = 64 - __builtin_clzll(10ull * (jb.d - kw) + 192);
        h = ipow2_mont<3>(E, bl, m, mp, mont1);
      }
      uint64_t q0, q1, q2;
      iquot192(h, m, mp, q0, q1, q2);
      add192(acc, q0, q1, q2);
      nterms += (m > 1);
    }
    k += kWave;
    m += mstep;
    e -= 10ull * kWave;
    if (carry_one) {
      const uint64_t t = (uint64_t)qq * mstep;  // < m
      if (one >= t) {
        one -= t;
      } else {
        one = one + m - t;
        qq -= 1;
      }
    }
  }
}

template <int KIND, int TPL>
__device__ __forceinline__ void tile_terms(const JobDev& jb, uint64_t kbase, int lane, U192& acc,
                                           uint64_t& nterms) {
  if constexpr (KIND == kSplitW3) tile_split<3, TPL>(jb, kbase, lane, acc, nterms);
  if constexpr (KIND == kSplitW2) tile_split<2, TPL>(jb, kbase, lane, acc, nterms);
  if constexpr (KIND == kSplitW1) tile_split<1, TPL>(jb, kbase, lane, acc, nterms);
  if constexpr (KIND == kPlainW3) tile_plain<kPlainW3, TPL>(jb, kbase, lane, acc, nterms);
  if constexpr (KIND == kPlainWide) tile_plain<kPlainWide, TPL>(jb, kbase, lane, acc, nterms);
}

// ------------------------------------------------------------------ tail terms

// Tail k = d + i, i = lane < 20: value 2^(-10i)/m (Bellard.cl:1164-1239).  Term =
// floor(2^(192-10i) / m), complemented when negative.  For m < 2^48 the quotient is the fp64 long
// division of the head terms' digits (first digit E mod 48 bits, then 48 bits each); larger m take
// bit-serial long division (exact for m < 2^63).
__device__ __forceinline__ U192 quot_pow2_f(int E, uint64_t mi) {
  const double m = (double)mi;
  const double minv = recip(m);
  const int a = E % 48 == 0 ? 48 : E % 48;
  const int c = (E - a) / 48;  // digits after the first; digit j has weight 2^(48 (c - j))
  U192 q = {0, 0, 0};
  double rem = 1.0;
  for (int j = 0; j <= c; ++j) {
    const double n = __builtin_amdgcn_ldexp(rem, j == 0 ? a : 48);
    // The estimate may be -1 only when n < m / 4 (the first digit, 2^a < m / 4): then the digit
    // is 0 and the remainder n.  Later remainders are nonzero residues of 2^(a + 48 j), so their
    // digits are at least 2^48 / m - 1 >= 0.
    const double qd = __builtin_fmax(__builtin_floor(__builtin_fma(n, minv, -0.25)), 0.0);
    rem = __builtin_fma(-qd, m, n);
    const uint64_t u = f2u64(qd);
    switch (c - j) {
      case 0: add192(q, u, 0, 0); break;
      case 1: add192(q, u << 48, u >> 16, 0); break;
      case 2: add192(q, 0, u << 32, u >> 32); break;
      default: add192(q, 0, 0, u << 16); break;
    }
  }
  if (rem >= m) add192(q, 1, 0, 0);  // lazy remainder in [m, 2m)
  return q;
}

__device__ __forceinline__ U192 quot_pow2_bits(int E, uint64_t m) {
  U192 q = {0, 0, 0};
  uint64_t rem = 1;  // numerator 2^E; its leading step leaves remainder 1
  for (int b = E - 1; b >= 0; --b) {
    rem <<= 1;
    if (rem >= m) {
      rem -= m;
      if (b >= 128)
        q.w2 |= 1ull << (b - 128);
      else if (b >= 64)
        q.w1 |= 1ull << (b - 64);
      else
        q.w0 |= 1ull << b;
    }
  }
  return q;
}

// One tail job on one wave: lane i < 20 takes k = d + i; the wave sum goes to copy `rep` of the
// job's slot.
__device__ __forceinline__ void tail_job(const TailDev tj, int lane, uint32_t replicas,
                                         uint32_t rep, unsigned long long* __restrict__ acc) {
  U192 q = {0, 0, 0};
  if (lane < kTailTerms) {
    const uint64_t k = tj.d + (uint64_t)lane;
    const uint64_t m = tj.den0 * k + tj.den1;
    const int E = 192 - 10 * lane;
    q = m < (1ull << 48) ? quot_pow2_f(E, m) : quot_pow2_bits(E, m);
    if ((tj.numesign + (uint32_t)(k & 1)) & 1u) q = {~q.w0, ~q.w1, ~q.w2};
  }
  wave_sum192(q);
  if (lane == 0) atomic_add192(acc + ((size_t)tj.slot * replicas + rep) * kAccWords, q);
}

// Clock samples of the launch (LaunchDev::clk_out): words at + 0, 1 at the start, at + 2, 3 when
// the wave leaves the tile loop.  Counter reads only, outside the term loop.
__device__ __forceinline__ void clock_sample(const LaunchDev* hdr, int at) {
  if (block_x() < (uint32_t)kClockXcds && thread_x() == 0 && hdr->clk_out) {
    uint64_t* clk = reinterpret_cast<uint64_t*>(hdr->clk_out) + 4 * block_x();
    clk[at] = __builtin_amdgcn_s_memtime();
    clk[at + 1] = __builtin_amdgcn_s_memrealtime();
  }
}

// Persistent head-term kernel; it also takes the tail jobs (wave w the tails w, w + nwaves, ...,
// before its tiles).  Flat tile f belongs to the last job with flat_begin <= f.  Wave w first
// takes the static range [w T, (w + 1) T) of T = static_tiles consecutive tiles (no atomics;
// small launches are all static), then chunks of consecutive tiles after the nwaves T static
// ones: chunk c (one atomic on a global counter per chunk) lies in the schedule's phase p with
// first_chunk[p] <= c < first_chunk[p + 1], whose chunks shrink from phase to phase (a guided
// schedule: few grabs while much work remains, small chunks for the final balance), so the launch
// balances itself whatever a tile costs.  A wave's tiles and chunks only increase, so its job
// index and phase only move forward.  Workgroups are waves_per_group waves (the grid's waves
// share no LDS and no barrier, so the size only affects how the hardware places them).
template <int KIND, int TPL>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_head(LaunchDev* __restrict__ hdr) {
  unsigned long long* __restrict__ acc = hdr_acc(hdr);
  const JobDev* __restrict__ jobs = hdr_jobs(hdr);
  const int njobs = (int)hdr->njobs;
  const uint64_t total_tiles = hdr->total_tiles;
  const uint32_t wpg = hdr->waves_per_group;
  const uint32_t nphases = hdr->nphases;
  // Wave counts and the static phase fit 32 bits (at most 64 static tiles per wave), which keeps
  // this bookkeeping in scalar registers.
  const uint32_t nwaves = grid_dim_x() * wpg;
  // readfirstlane: the compiler does not know the work-item index / 64 is uniform, and a
  // divergent tile index would move the job table loads and all per-tile values into vector
  // registers.
  const uint32_t wave_id = block_x() * wpg + __builtin_amdgcn_readfirstlane(thread_x() / kWave);
  const uint32_t static_tiles = (uint32_t)hdr->static_tiles;
  const uint64_t dyn0 = static_tiles * nwaves;  // first tile of the dynamic phase
  const int lane = (int)(thread_x() & (kWave - 1));
  // Slot s has `replicas` copies; wave w adds into copy w mod replicas, which spreads the atomics
  // of many short flushes over more addresses (atomics on one address serialise).
  const uint32_t replicas = hdr->replicas;
  const uint32_t rep = wave_id & (replicas - 1);
  clock_sample(hdr, 0);
  {
    const TailDev* __restrict__ tails = hdr_tails(hdr);
    for (uint32_t t = wave_id; t < hdr->ntails; t += nwaves)
      tail_job(tails[t], lane, replicas, rep, acc);
  }
  int cur = -1;
  JobDev jb{};
  U192 sum = {0, 0, 0};
  uint64_t nterms = 0;
  auto tile = [&](uint64_t f) {
    int lo = (cur < 0) ? 0 : cur, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].flat_begin <= f)
        lo = mid;
      else
        hi = mid - 1;
    }
    if (lo != cur) {
      if (cur >= 0)
        flush_job(sum, nterms, jb.numesign, lane,
                  acc + ((size_t)jb.slot * replicas + rep) * kAccWords);
      cur = lo;
      jb = jobs[lo];
      sum = {0, 0, 0};
      nterms = 0;
    }
    const uint64_t kbase = (jb.tile_first + (f - jb.flat_begin)) * tile_k(TPL);
    tile_terms<KIND, TPL>(jb, kbase, lane, sum, nterms);
  };
  // One loop and one call site of tile(), so the term code is emitted once (two copies cost
  // registers and occupancy).  All branches here are wave-uniform.
  uint64_t f = wave_id * static_tiles;  // < 2^32: 64 static tiles per wave at most
  uint64_t f_end = f + static_tiles < total_tiles ? f + static_tiles : total_tiles;
  uint32_t phase = 0;
  for (;;) {
    if (f >= f_end) {
      if (dyn0 >= total_tiles || nphases == 0) break;
      unsigned long long got = 0;
      if (lane == 0) got = atomicAdd(&hdr->counter, 1ull);
      const uint32_t c = __builtin_amdgcn_readfirstlane((uint32_t)got);
      while (phase + 1 < nphases && c >= hdr->phases[phase + 1].first_chunk) ++phase;
      const uint32_t chunk = hdr->phases[phase].chunk;
      f = dyn0 + hdr->phases[phase].start +
          (uint64_t)(c - hdr->phases[phase].first_chunk) * (uint64_t)chunk;
      if (f >= total_tiles) break;
      f_end = f + chunk < total_tiles ? f + chunk : total_tiles;
    }
    tile(f);
    ++f;
  }
  clock_sample(hdr, 2);
  if (cur >= 0)
    flush_job(sum, nterms, jb.numesign, lane, acc + ((size_t)jb.slot * replicas + rep) * kAccWords);
}

// ------------------------------------------------------------------ combine

__global__ void k_zero(unsigned long long* p, uint64_t n) {
  for (uint64_t i = block_x() * (uint64_t)block_dim_x() + thread_x(); i < n;
       i += (uint64_t)grid_dim_x() * block_dim_x())
    p[i] = 0;
}

// Carry-propagate six 32-bit limb sums into a 192-bit value.
__device__ __forceinline__ U192 limbs_value(const uint64_t (&s)[6]) {
  uint64_t limb[6];
  uint64_t c = 0;
  for (int i = 0; i < 6; ++i) {
    const uint64_t t = s[i] + c;  // s[i] < 2^63 while a slot takes fewer than 2^31 additions
    limb[i] = t & 0xffffffffull;
    c = t >> 32;
  }
  return {limb[0] | (limb[1] << 32), limb[2] | (limb[3] << 32), limb[4] | (limb[5] << 32)};
}

__device__ __forceinline__ U192 shl192(U192 v, int sh) {
  if (sh == 0) return v;
  return {v.w0 << sh, (v.w1 << sh) | (v.w0 >> (64 - sh)), (v.w2 << sh) | (v.w1 >> (64 - sh))};
}

__device__ __forceinline__ void store192(uint64_t* p, const U192& v) {  // MSW first
  p[0] = v.w2;
  p[1] = v.w1;
  p[2] = v.w0;
}

// Sum7_192-style combine (Bellard.cl:1359-1398): per group, every slot's replicas summed (lane r
// loads copy r), carried into 192 bits, shifted by log2 c_j and added to the group's total.  One
// block of kCombineWaves waves per group (grid-stride over the groups); wave w takes the group's
// slots w, w + kCombineWaves, ..., and the waves' partial totals meet in LDS.
__global__ void __launch_bounds__(kCombineWaves * kWave) k_combine(LaunchDev* __restrict__ hdr) {
  __shared__ uint64_t part[kCombineWaves][3];
  const unsigned long long* __restrict__ acc = hdr_acc(hdr);
  const GroupDev* __restrict__ groups = hdr_groups(hdr);
  const uint8_t* __restrict__ shift = hdr_shifts(hdr);
  uint64_t* dst = reinterpret_cast<uint64_t*>(hdr->dst);
  uint64_t* slot_out = reinterpret_cast<uint64_t*>(hdr->slot_out);
  const uint32_t replicas = hdr->replicas;
  const int lane = (int)(thread_x() & (kWave - 1));
  const uint32_t wave = thread_x() / kWave;
  for (uint32_t g = block_x(); g < hdr->ngroups; g += grid_dim_x()) {
    const GroupDev gr = groups[g];
    U192 total = {0, 0, 0};
    for (uint32_t i = wave; i < gr.count; i += kCombineWaves) {
      const uint32_t s = gr.first + i;
      const unsigned long long* base = acc + (size_t)s * replicas * kAccWords;
      uint64_t v[6];
#pragma unroll
      for (int k = 0; k < 6; ++k)
        v[k] = (uint32_t)lane < replicas ? base[(size_t)lane * kAccWords + k] : 0;
      if (replicas > 1) {
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1)
#pragma unroll
          for (int k = 0; k < 6; ++k) v[k] += shfl_xor64(v[k], mask);
      }
      const U192 val = shl192(limbs_value(v), shift[s]);
      if (lane == 0) {
        if (slot_out) store192(slot_out + 3 * (size_t)s, val);
        if (hdr->mirror_slot && s == hdr->mirror_index)
          store192(reinterpret_cast<uint64_t*>(hdr->mirror_slot), val);
      }
      add192(total, val.w0, val.w1, val.w2);
    }
    if (lane == 0) {
      part[wave][0] = total.w0;
      part[wave][1] = total.w1;
      part[wave][2] = total.w2;
    }
    __syncthreads();
    if (thread_x() == 0) {
      U192 sum = {0, 0, 0};
      if (hdr->accumulate) sum = {dst[3 * g + 2], dst[3 * g + 1], dst[3 * g + 0]};
      for (int w = 0; w < kCombineWaves; ++w) add192(sum, part[w][0], part[w][1], part[w][2]);
      store192(dst + 3 * (size_t)g, sum);
      if (hdr->mirror_total && g == 0)
        store192(reinterpret_cast<uint64_t*>(hdr->mirror_total), sum);
    }
    __syncthreads();
  }
}

// Start-up check of the FP64 output modifier that the doubling squares rely on: one lane runs
// fsq_bits on 3 mod 1000003 with the exponent bit set (2 * 3^2 = 18) and clear (3^2 = 9).  A
// device build with the IEEE mode bit set ignores the modifier and returns 9 for both.
__global__ void k_selfcheck(uint64_t* out) {
  if (thread_x() != 0) return;
  const double m = 1000003.0;
  const double minv = recip(m);
  out[0] = (uint64_t)(int64_t)fsq_bits(3.0, 2ull, 1u, m, minv);
  out[1] = (uint64_t)(int64_t)fsq_bits(3.0, 1ull, 1u, m, minv);
}

__global__ void k_sum7(uint64_t* ans, uint64_t* partial, int shift) {
  if (thread_x() != 0 || block_x() != 0) return;
  const U192 p = shl192({partial[2], partial[1], partial[0]}, shift);
  partial[0] = p.w2;
  partial[1] = p.w1;
  partial[2] = p.w0;
  U192 a = {ans[2], ans[1], ans[0]};
  add192(a, p.w0, p.w1, p.w2);
  ans[0] = a.w2;
  ans[1] = a.w1;
  ans[2] = a.w0;
}

// ------------------------------------------------------------------ launch wrappers

template <int KIND, int TPL>
static int resident_blocks(int block) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_head<KIND, TPL>, block, 0) != hipSuccess)
    n = 0;
  return n;
}

template <int KIND>
static int resident_blocks_tpl(int tpl_index, int block) {
  switch (tpl_index) {
    case 0: return resident_blocks<KIND, kTplChoices[0]>(block);
    case 1: return resident_blocks<KIND, kTplChoices[1]>(block);
    case 2: return resident_blocks<KIND, kTplChoices[2]>(block);
    default: return 0;
  }
}

int head_blocks_per_cu(int kind, int tpl_index, int waves_per_group) {
  const int block = kWave * waves_per_group;
  switch (kind) {
    case kSplitW3: return resident_blocks_tpl<kSplitW3>(tpl_index, block);
    case kSplitW2: return resident_blocks_tpl<kSplitW2>(tpl_index, block);
    case kSplitW1: return resident_blocks_tpl<kSplitW1>(tpl_index, block);
    case kPlainW3: return resident_blocks_tpl<kPlainW3>(tpl_index, block);
    case kPlainWide: return resident_blocks_tpl<kPlainWide>(tpl_index, block);
    default: return 0;
  }
}

template <int KIND>
static hipError_t launch_head_tpl(int tpl_index, LaunchDev* hdr, int grid, int block,
                                  hipStream_t stream) {
  switch (tpl_index) {
    case 0:
      hipLaunchKernelGGL((k_head<KIND, kTplChoices[0]>), dim3(grid), dim3(block), 0, stream, hdr);
      break;
    case 1:
      hipLaunchKernelGGL((k_head<KIND, kTplChoices[1]>), dim3(grid), dim3(block), 0, stream, hdr);
      break;
    case 2:
      hipLaunchKernelGGL((k_head<KIND, kTplChoices[2]>), dim3(grid), dim3(block), 0, stream, hdr);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_head(int kind, int tpl_index, LaunchDev* hdr, int grid, int waves_per_group,
                       hipStream_t stream) {
  if (waves_per_group < 1 || waves_per_group > kMaxBlock / kWave) return hipErrorInvalidValue;
  const int block = kWave * waves_per_group;
  switch (kind) {
    case kSplitW3: return launch_head_tpl<kSplitW3>(tpl_index, hdr, grid, block, stream);
    case kSplitW2: return launch_head_tpl<kSplitW2>(tpl_index, hdr, grid, block, stream);
    case kSplitW1: return launch_head_tpl<kSplitW1>(tpl_index, hdr, grid, block, stream);
    case kPlainW3: return launch_head_tpl<kPlainW3>(tpl_index, hdr, grid, block, stream);
    case kPlainWide: return launch_head_tpl<kPlainWide>(tpl_index, hdr, grid, block, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_combine(LaunchDev* hdr, hipStream_t stream) {
  hipLaunchKernelGGL(k_combine, dim3(kCombineBlocks), dim3(kCombineWaves * kWave), 0, stream,
                     hdr);
  return hipGetLastError();
}

hipError_t launch_zero(unsigned long long* p, uint64_t nwords, hipStream_t stream) {
  if (nwords == 0) return hipSuccess;
  const uint64_t blocks = (nwords + 255) / 256;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, stream,
                     p, nwords);
  return hipGetLastError();
}

hipError_t launch_sum7(uint64_t* ans, uint64_t* partial, int shift, hipStream_t stream) {
  hipLaunchKernelGGL(k_sum7, dim3(1), dim3(64), 0, stream, ans, partial, shift);
  return hipGetLastError();
}

hipError_t launch_selfcheck(uint64_t* out2, hipStream_t stream) {
  hipLaunchKernelGGL(k_selfcheck, dim3(1), dim3(64), 0, stream, out2);
  return hipGetLastError();
}

}  // namespace clb

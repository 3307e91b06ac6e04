// clb_internal.hpp -- structures shared by the host code and the kernels of libclbellard.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clb {

constexpr int kWave = 64;
constexpr int kTermsPerLane = 16;                          // terms per lane per tile
constexpr uint64_t kTileK = (uint64_t)kWave * kTermsPerLane;  // consecutive k per tile (even)
constexpr int kBlock = 256;                                // threads per workgroup
constexpr int kAccWords = 8;                               // u64 words per accumulator slot
constexpr int kTailTerms = 20;                             // k = d .. d+19 (Bellard.cl:1164-1239)

// Term kernel variants.
enum Kind : int {
  kSplitW3 = 0,  // fp64 engine, 3-bit windows: m < 2^42
  kSplitW2 = 1,  // fp64 engine, 2-bit windows: m < 2^46
  kSplitW1 = 2,  // fp64 engine, 1-bit windows: m < 2^48
  kPlainW3 = 3,  // int engine, lazy residues, 3-bit windows: m < 2^55
  kPlainWide = 4 // int engine, reduced residues, 1-bit windows: m < 2^63
};

// One head-term job in the flat tile space.  Tiles are aligned to absolute multiples of kTileK,
// so lane L of every tile handles k = L (mod 64) and keeps one sign for the whole job.
struct JobDev {
  uint64_t d;           // loop count
  uint64_t k0, k1;      // head range [k0, k1), k1 <= d, k0 < k1
  uint64_t tile_first;  // k0 / kTileK
  uint64_t flat_begin;  // first flat tile index of the job (prefix sum, ascending)
  uint64_t den0, den1;  // modulus m = den0 * k + den1
  uint32_t numesign;    // 1: the term of even k is negative (KernelFuncs.as:10)
  uint32_t slot;        // accumulator slot
};

// The 20 tail terms of one formula at loop count d (k = d .. d+19).
struct TailDev {
  uint64_t d, den0, den1;
  uint32_t numesign;
  uint32_t slot;
};

// One result: the sum over slots [first, first + count) of (slot value << shift[slot]) mod 2^192.
struct GroupDev {
  uint32_t first;
  uint32_t count;
};

// Workgroups of k_head<kind> that fit on one compute unit at once (0 if unknown).
int head_blocks_per_cu(int kind);

// Kernel launch wrappers (clb_kernels.hip).  All asynchronous on `stream`.
// counter: one zeroed u64 the waves take chunks of `chunk` tiles from.
hipError_t launch_head(int kind, const JobDev* jobs, int njobs, uint64_t total_tiles,
                       uint32_t chunk, unsigned long long* counter, unsigned long long* acc,
                       int grid, hipStream_t stream);
hipError_t launch_tail(const TailDev* tails, int ntails, unsigned long long* acc,
                       hipStream_t stream);
hipError_t launch_zero(unsigned long long* p, uint64_t nwords, hipStream_t stream);
// out[g] (3 words, most significant first) = sum of shifted slots of group g; accumulate adds
// into out instead of overwriting.  slot_out (may be null) receives each shifted slot value.
hipError_t launch_finalize(const unsigned long long* acc, const uint8_t* shift, int nslots,
                           const GroupDev* groups, int ngroups, uint64_t* out, int accumulate,
                           uint64_t* slot_out, hipStream_t stream);
// Sum7_192 (Bellard.cl:1359-1398): partial <<= shift (written back), ans += partial.
hipError_t launch_sum7(uint64_t* ans, uint64_t* partial, int shift, hipStream_t stream);

}  // namespace clb

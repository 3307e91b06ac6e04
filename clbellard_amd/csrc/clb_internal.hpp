// clb_internal.hpp -- structures shared by the host code and the kernels of libclbellard.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

namespace clb {

constexpr int kWave = 64;
// Terms per lane per tile (TPL): a tile is kWave * TPL consecutive k, lane L taking k = base + L +
// 64 i, i < TPL.  16 is the most the fp64 engine's digit sums hold exactly; launches too small to
// fill the GPU with 16-term tiles use 4 or 1 so that more waves run side by side.
constexpr int kMaxTermsPerLane = 16;
constexpr int kNumTpl = 3;
constexpr int kTplChoices[kNumTpl] = {16, 4, 1};
__host__ __device__ constexpr uint64_t tile_k(int tpl) { return (uint64_t)kWave * (uint64_t)tpl; }
constexpr int kBlock = 256;     // threads per workgroup (k_head's default; the other kernels')
constexpr int kMaxBlock = 512;  // k_head's largest workgroup (8 waves)
constexpr int kMaxPhases = 24;  // phases of the tile queue's guided schedule
constexpr int kAccWords = 8;                               // u64 words per accumulator slot
constexpr int kTailTerms = 20;                             // k = d .. d+19 (Bellard.cl:1164-1239)

// Term kernel variants.
enum Kind : int {
  kSplitW3 = 0,  // fp64 engine, 3-bit windows: m < 2^42
  kSplitW2 = 1,  // fp64 engine, 2-bit windows: m < 2^46
  kSplitW1 = 2,  // fp64 engine, 1-bit windows: m < 2^48
  kPlainW3 = 3,   // int engine, lazy residues, 3-bit windows: m < 2^55
  kPlainWide = 4, // int engine, reduced residues, 1-bit windows: m < 2^63
  kNumKinds = 5
};

// One head-term job in the flat tile space.  Tiles are aligned to absolute multiples of the tile
// size (a multiple of 64), so lane L of every tile handles k = L (mod 64) and keeps one sign for
// the whole job.
struct JobDev {
  uint64_t d;           // loop count
  uint64_t k0, k1;      // head range [k0, k1), k1 <= d, k0 < k1
  uint64_t tile_first;  // k0 / tile size
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

// Reference replay: one launch of the reference's Sglobal32mtg_192 / Sglobal64mtg_192
// (Bellard.cl:771-823, :992-1084) with `gsize` work-items.  Work-item idx < min(gsize, k_max -
// offset) sums the terms k = offset + idx + i gsize < k_max in double-double and converts once;
// its tiles are 64 consecutive work-items, in the flat replay tile space from flat_begin.
struct ReplayJobDev {
  uint64_t d;
  uint64_t offset, k_max;
  uint64_t gsize;
  uint64_t flat_begin;
  uint64_t den0, den1;
  uint32_t numesign;
  uint32_t slot;
};

// One result: the sum over slots [first, first + count) of (slot value << shift[slot]) mod 2^192.
struct GroupDev {
  uint32_t first;
  uint32_t count;
};

// One phase of the dynamic tile schedule: chunks first_chunk, first_chunk + 1, ... of `chunk`
// consecutive tiles each, starting at tile `start` (relative to the end of the static phase); the
// last phase runs to the end of the tile space.  Chunk indices stay below 2^31 (the host sizes the
// chunks so), so the kernel compares and multiplies them on the scalar unit in 32 bits.
struct QueuePhase {
  uint64_t start;
  uint32_t first_chunk;
  uint32_t chunk;
};

// Per-call header at offset 0 of the call's block.  The block holds, at the byte offsets below,
// the job table, the tail table, the groups, the shifts, the accumulator (zeroed by the copy that
// brings the block in) and room for the results.  The kernels read everything from the block, so
// their only argument is its address, and one captured HIP graph serves every call while the
// block is not reallocated.  The host copies the block before each launch sequence, which also
// resets the tile queue.
struct LaunchDev {
  unsigned long long counter;  // tile queue: chunks handed out in the dynamic phase
  uint64_t total_tiles;        // flat tiles of all jobs
  uint64_t static_tiles;       // static phase: wave w takes tiles [w T, (w + 1) T), T = this
  uint64_t dst;                // device address of the group results (3 words each, MSW first)
  uint64_t slot_out;           // 0, or the device address of every slot's shifted value
  uint64_t mirror_total;       // 0, or a device address that also receives group 0's result
  uint64_t mirror_slot;        // 0, or a device address that also receives slot mirror_index
  uint64_t clk_out;            // 0, or the device address of kClockXcds x 4 words: k_head's clock
                               // samples (below)
  uint32_t mirror_index;
  uint32_t accumulate;         // 1: dst += result (the AddAns-style accumulation)
  uint32_t njobs;              // entries of the job table
  uint32_t waves_per_group;    // k_head's workgroup size in waves (1, 2, 4 or 8)
  uint32_t ntails;             // entries of the tail table
  uint32_t ngroups;            // results (groups of slots)
  uint32_t replicas;           // copies per accumulator slot (a power of 2, at most 64)
  uint32_t nslots;             // accumulator slots
  uint32_t off_jobs, off_tails, off_groups, off_shifts, off_acc, off_out;  // byte offsets
  // Reference replay (0 entries unless the call replays the reference's kernels): its launches'
  // work-item tiles from their own queue, and its tails (Sglobal_after_192, one lane each).
  unsigned long long rcounter;  // replay tiles handed out
  uint64_t rtotal_tiles;
  uint32_t nrjobs, nrtails, rchunk;
  uint32_t off_rjobs, off_rtails;
  uint32_t nphases;                   // phases of the dynamic tile schedule (0: all static)
  QueuePhase phases[kMaxPhases];
};

__host__ __device__ inline const uint8_t* hdr_at(const LaunchDev* h, uint32_t off) {
  return reinterpret_cast<const uint8_t*>(h) + off;
}
__host__ __device__ inline const JobDev* hdr_jobs(const LaunchDev* h) {
  return reinterpret_cast<const JobDev*>(hdr_at(h, h->off_jobs));
}
__host__ __device__ inline const TailDev* hdr_tails(const LaunchDev* h) {
  return reinterpret_cast<const TailDev*>(hdr_at(h, h->off_tails));
}
__host__ __device__ inline const GroupDev* hdr_groups(const LaunchDev* h) {
  return reinterpret_cast<const GroupDev*>(hdr_at(h, h->off_groups));
}
__host__ __device__ inline const uint8_t* hdr_shifts(const LaunchDev* h) {
  return hdr_at(h, h->off_shifts);
}
__host__ __device__ inline const ReplayJobDev* hdr_rjobs(const LaunchDev* h) {
  return reinterpret_cast<const ReplayJobDev*>(hdr_at(h, h->off_rjobs));
}
__host__ __device__ inline const TailDev* hdr_rtails(const LaunchDev* h) {
  return reinterpret_cast<const TailDev*>(hdr_at(h, h->off_rtails));
}
__host__ __device__ inline unsigned long long* hdr_acc(LaunchDev* h) {
  return reinterpret_cast<unsigned long long*>(reinterpret_cast<uint8_t*>(h) + h->off_acc);
}

// k_head's clock samples: the first wave of workgroup x < kClockXcds (workgroups are dispatched
// round-robin over the 8 XCDs, so one per XCD) writes, at clk_out[4 x ..], the shader-clock
// counter (s_memtime) and the 100 MHz real-time counter (s_memrealtime) when it starts and when
// it leaves the tile loop: the clock that XCD ran the launch at.
constexpr int kClockXcds = 8;

constexpr int kCombineBlocks = 256;  // k_combine grid (one block per group, grid-stride)
constexpr int kCombineWaves = 8;     // k_combine block: waves sharing a group's slots

int head_blocks_per_cu(int kind, int tpl_index, int waves_per_group);

// The launch sequence of one call: the head and tail terms (persistent grid of `grid`
// workgroups), then the combine (slot replicas folded, shifted, summed per group).
hipError_t launch_head(int kind, int tpl_index, LaunchDev* hdr, int grid, int waves_per_group,
                       hipStream_t stream);
hipError_t launch_combine(LaunchDev* hdr, hipStream_t stream);
hipError_t launch_zero(unsigned long long* p, uint64_t nwords, hipStream_t stream);
hipError_t launch_sum7(uint64_t* ans, uint64_t* partial, int shift, hipStream_t stream);
// Writes 18 and 9 to out2[0..1] when the device code applies FP64 output modifiers.
hipError_t launch_selfcheck(uint64_t* out2, hipStream_t stream);
hipError_t launch_replay_selfcheck(uint64_t* out2, hipStream_t stream);

// Reference replay (clb_replay.hip): persistent grid over the replay tiles and tails of `hdr`.
// Reference replay of the 106-bit mode (precision 0): per BatchCalc job, the reference's 65536
// double-double group accumulators (bigSumMem, Bellard_192.hsp:409-411) in device memory.
constexpr uint32_t kRefGroups = 65536;  // BIGSUM_SIZE, Bellard_192.hsp:5
struct R106JobDev {
  uint64_t d, k0, k1;    // BatchCalc head range after the clamp to d (k0 >= k1: no head terms)
  uint64_t den0, den1;
  uint64_t bigsum;       // device address of the job's kRefGroups accumulators (hi, lo doubles)
  uint64_t part;         // device address receiving bigSum[0] after Ssum (hi, lo)
  uint32_t numesign, tail;
  uint32_t single;       // 1: [k0, k1) is one launch of one kernel (Sglobal32mtg if m(k0) < 2^32,
                         // else Sglobal64mtg), not BatchCalc's chunks: the operator clb_sglobal
  uint32_t ssum;         // 1: the two Ssum trees after the tail (BatchCalc; clb_ssum)
};
// kind: kSplitW3/W2/W1 (fp64 residues) or kPlainW3 (integer residues), by the largest modulus.
// finish: 1 to run the tail / Ssum kernel after the group accumulators (jobs with tail or ssum).
int replay_blocks_per_cu(int kind);
hipError_t launch_replay(int kind, LaunchDev* hdr, int grid, hipStream_t stream);
// 106-bit mode: the group accumulators of jobs[0..njobs) (kind as above), then per job the tail
// and the two Ssum trees (writing each job's part), then AddAns per result group: ans[g] (in: the
// starting answer, out: the answer) and slot_val[s] (out: part after Sum7's update), where group g
// holds slots [first[g], first[g] + count[g]) with numerators 2^shift[s].  Doubles in (hi, lo).
hipError_t launch_replay106(int kind, const R106JobDev* jobs, uint32_t njobs, uint64_t gsize,
                            uint32_t ngroups_max, hipStream_t stream);
hipError_t launch_replay106_sum7(const double* parts, const uint32_t* first, const uint32_t* count,
                                 const uint8_t* shift, double* ans, double* slot_val,
                                 uint32_t ngroups, hipStream_t stream);
constexpr uint64_t kRefGlobalWorkSize = 65536ull * 256ull;  // Bellard_192.hsp:5-7

// The no-allocation entry point (clb_noalloc.hip): a fixed device array per GPU and a pinned
// staging buffer, each kNoAllocBytes, set up once per device (with that device current).  A call
// of at most kNoAllocMaxJobs jobs always fits: its block is under 1 KiB of header, 64 B per job,
// 32 B per tail, 8 + 1 + 24 B per result and 64 B per accumulator copy, and the copies number at
// most max(jobs, waves / 32).
constexpr size_t kNoAllocBytes = 1u << 20;
constexpr uint32_t kNoAllocMaxJobs = 4096;
hipError_t noalloc_setup(int device);
// The device's array address, its staging buffer and the lock a call holds; false before setup.
bool noalloc_area(int device, uint8_t** dev, uint8_t** host, std::mutex** call);

}  // namespace clb

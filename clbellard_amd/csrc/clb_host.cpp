// clb_host.cpp -- host side of libclbellard.so: the C ABI declared in include/clbellard.h.
//
// This replaces the reference's HSP host (Bellard_192.hsp, KernelFuncs.as) and its OpenCL plugin
// calls with one C++ library over the HIP runtime.  Every computation of terms happens in the
// kernels of clb_kernels.hip; this file builds job tables, moves a few hundred bytes and runs work
// mode.  The host-only entry points (formats, worklists, work files, the summer) are in
// clb_format.cpp.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clbellard.h"
#include "clb_hostlib.hpp"
#include "clb_internal.hpp"

using namespace clb;

namespace {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(CLB_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                  __FILE__, __LINE__);                                                   \
  } while (0)

// Head-term job before tiling.
struct HostJob {
  uint64_t d, k0, k1;  // k0 < k1 <= d, or k0 == k1 for no head terms
  uint64_t den0, den1;
  uint32_t numesign;
  uint32_t slot;
  bool tail;
  bool single = false;  // one kernel launch over [k0, k1) (clb_sglobal), not BatchCalc's chunks
  bool ssum_op = false;  // clb_ssum in the 106-bit-mode replay: only Ssum's two trees
};

// Device buffer that grows geometrically, so that its address (which a captured graph records)
// changes rarely.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  int reserve(size_t n) {
    if (n <= cap) return CLB_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    size_t want = n < 64 ? 64 : n;
    if (want < 2 * cap) want = 2 * cap;
    cap = 0;
    hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
    if (e != hipSuccess) return fail(CLB_ENOMEM, "hipMalloc(%zu bytes): %s", want * sizeof(T),
                                     hipGetErrorString(e));
    cap = want;
    return CLB_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Pinned host staging (hipHostMalloc), so that the per-call copies are single DMA transfers.
struct HostBuf {
  uint8_t* p = nullptr;
  size_t cap = 0;
  int reserve(size_t n) {
    if (n <= cap) return CLB_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    size_t want = n < 4096 ? 4096 : n;
    if (want < 2 * cap) want = 2 * cap;
    cap = 0;
    hipError_t e = hipHostMalloc((void**)&p, want, hipHostMallocDefault);
    if (e != hipSuccess)
      return fail(CLB_ENOMEM, "hipHostMalloc(%zu bytes): %s", want, hipGetErrorString(e));
    cap = want;
    return CLB_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace

struct clb_ctx {
  int device = 0;
  int precision = CLB_PREC_U192;
  int kernel = CLB_KERNEL_AUTO;
  int blocks_per_cu = 0;
  int num_cus = 0;
  char device_name[256] = {0};
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // around the head kernel (or the replayed graph)
  uint64_t* partial = nullptr;  // bigSumMem[0..2] (most significant word first)
  uint64_t* ans = nullptr;      // dev_ans[0..2]
  DevBuf<uint8_t> block;        // the call's block: header, tables, accumulator, results
  HostBuf block_host;           // its pinned staging copy (results land here too)
  float head_ms = 0.f, total_ms = 0.f;
  // k_head's clock samples (kClockXcds x 4 words), and whether the last call's head kernel wrote
  // them
  uint64_t* clk = nullptr;
  bool clk_valid = false;
  // resident k_head workgroups per CU, per kind, tile length and workgroup size (1, 2, 4, 8 waves)
  int resident[kNumKinds][kNumTpl][4] = {};
  // Tile schedule knobs (clb_set_schedule): k_head's workgroup size in waves, the queue chunk
  // (the last phase's, for the guided schedule), the chunks per wave in each phase of the guided
  // schedule (0: fixed chunks, the default), the most copies per accumulator slot.
  int wg_waves = 4;
  int min_chunk = 256;
  int chunks_per_wave = 0;
  int max_replicas = 64;
  int tpl_index = -1;            // index into kTplChoices, or -1: chosen per launch
  int min_tiles_per_wave = 0;    // launch-size threshold of that choice (0: a tile per 8 waves)
  // Graph mode: the kernel sequence of a call, captured once per buffer layout and kernel choice
  // (every argument the kernels see), replayed afterwards.  Counts travel in hdr.
  bool use_graph = false;
  // Reference replay: 0, or the work-items per replayed launch of the reference's kernels.
  uint64_t replay_gws = 0;
  // Work mode: bounds of one launch of a worklist's jobs (0 = no bound), and how many launches one
  // clb_run_worklist call may make (0 = until the list is done).  Files are written per launch.
  uint32_t wl_max_jobs = 0;
  double wl_max_seconds = 60.0;
  uint32_t wl_max_launches = 0;
  // Estimated modular squares per second of the term kernels, refined by every worklist launch.
  double wl_rate = 5e12;
  int replay_resident[kNumKinds] = {-1, -1, -1, -1, -1};  // resident k_replay workgroups per CU, per kind
  // 106-bit-mode replay: the call's tables and results, the jobs' group accumulators, and the
  // context's own bigSumMem (65536 double-doubles) for BatchCalc / AddAns at the operator level.
  DevBuf<uint8_t> r106_block;
  HostBuf r106_host;
  DevBuf<uint8_t> r106_bigsum;
  DevBuf<uint8_t> r106_state;
  struct Graph {
    uintptr_t key[4];
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  hipError_t noalloc_err = hipSuccess;  // set-up result of this device's no-allocation area
};

namespace {

int set_device(clb_ctx* ctx) {
  HIP_TRY(hipSetDevice(ctx->device));
  return CLB_OK;
}

// Term kernel for the largest modulus of a launch (bounds in clb_device.hpp / clb_internal.hpp).
int choose_kind(int pref, uint64_t max_m, int* kind) {
  const bool split_ok = max_m < (1ull << 48);
  if (pref == CLB_KERNEL_SPLIT && !split_ok)
    return fail(CLB_ERANGE, "split kernel needs moduli < 2^48 (largest is %llu)",
                (unsigned long long)max_m);
  if (pref == CLB_KERNEL_SPLIT || (pref == CLB_KERNEL_AUTO && split_ok)) {
    *kind = max_m < (1ull << 42) ? kSplitW3 : max_m < (1ull << 46) ? kSplitW2 : kSplitW1;
    return CLB_OK;
  }
  if (max_m >= (1ull << 63))
    return fail(CLB_ERANGE, "modulus %llu exceeds 2^63", (unsigned long long)max_m);
  *kind = max_m < (1ull << 55) ? kPlainW3 : kPlainWide;
  return CLB_OK;
}

// Persistent grid: every workgroup resident at once (one wave of blocks), unless overridden.
int wg_index(int waves) { return waves == 1 ? 0 : waves == 2 ? 1 : waves == 4 ? 2 : 3; }

int blocks_per_cu(clb_ctx* ctx, int kind, int tpl) {
  if (ctx->blocks_per_cu > 0) return ctx->blocks_per_cu;
  int& res = ctx->resident[kind][tpl][wg_index(ctx->wg_waves)];
  if (res == 0) {
    const int r = head_blocks_per_cu(kind, tpl, ctx->wg_waves);
    res = r > 0 ? r : 32 / ctx->wg_waves;
  }
  return res;
}

// The captured launch sequence (head with tails, combine) for a kernel choice and block address,
// captured on first use.  Kernels read everything else from the block, so a graph stays valid
// while the block is not reallocated; at most 8 are kept.
int graph_for(clb_ctx* ctx, int kind, int tpl, int grid, hipGraphExec_t* exec) {
  LaunchDev* hdr = reinterpret_cast<LaunchDev*>(ctx->block.p);
  const uintptr_t key[4] = {(uintptr_t)kind | ((uintptr_t)ctx->wg_waves << 8), (uintptr_t)tpl,
                            (uintptr_t)grid, (uintptr_t)hdr};
  for (const auto& g : ctx->graphs)
    if (memcmp(g.key, key, sizeof key) == 0) {
      *exec = g.exec;
      return CLB_OK;
    }
  hipStream_t s = ctx->stream;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_head(kind, tpl, hdr, grid, ctx->wg_waves, s);
  if (e == hipSuccess) e = launch_combine(hdr, s);
  hipGraph_t graph = nullptr;
  const hipError_t e_end = hipStreamEndCapture(s, &graph);
  if (e == hipSuccess) e = e_end;
  hipGraphExec_t ex = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(CLB_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
  if (ctx->graphs.size() >= 8) {
    (void)hipGraphExecDestroy(ctx->graphs.front().exec);
    ctx->graphs.erase(ctx->graphs.begin());
  }
  clb_ctx::Graph g;
  memcpy(g.key, key, sizeof key);
  g.exec = ex;
  ctx->graphs.push_back(g);
  *exec = ex;
  return CLB_OK;
}

// Launch plan of an exact-mode call.
struct LaunchPlan {
  int tpl = 0;                // index into kTplChoices
  int wg_waves = 4;           // workgroup size in waves
  uint64_t grid = 0;          // persistent workgroups
  uint64_t nwaves = 0;
  uint64_t flat = 0;          // tiles of the call's head jobs
  uint64_t static_tiles = 0;  // tiles per wave of the static phase (0: none)
  uint32_t nphases = 0;       // phases of the dynamic schedule (0: all static)
  QueuePhase phases[kMaxPhases] = {};
  uint32_t replicas = 1;      // copies per accumulator slot
};

// Terms per lane and grid: the longest tiles that still give every resident wave
// ctx->min_tiles_per_wave tiles, or by default one tile to every 8th wave (only the smallest
// launches, e.g. N = 10^5, take shorter tiles: 4 terms per lane measured 11% faster there).
// flat_of(tile_k) gives the call's tile count for a tile length.  Then the tile schedule: up to 64
// tiles per wave, each wave takes one static range of consecutive tiles (no atomics; atomics on
// one address serialise at ~35 ns each); larger launches take all tiles from the queue in fixed
// chunks, at least 8 per wave and at most min_chunk = 256 tiles (~15 ms of one wave's work at
// N = 10^12, the granularity of the final balance).  With chunks_per_wave > 0 the queue follows a
// guided schedule instead: each phase hands out about half of the remaining tiles in
// chunks_per_wave chunks per wave, so chunks shrink geometrically down to min_chunk, and the queue
// takes about nwaves chunks_per_wave log2(tiles / nwaves) grabs instead of tiles / min_chunk;
// measured 2.7-11% slower at N = 10^11 (1, 2, 4 chunks per wave) than fixed chunks, so it is not
// the default.  And the copies per accumulator slot: enough that the waves' flushes spread over
// about nwaves / 16 addresses per limb (a power of 2, at most max_replicas), so short launches do
// not queue on one slot.
template <class FlatOf>
void plan_launch(clb_ctx* ctx, int kind, size_t nslots, FlatOf flat_of, LaunchPlan* p) {
  p->wg_waves = ctx->wg_waves;
  int tpl = ctx->tpl_index;
  for (int t = (tpl >= 0 ? tpl : 0); t < kNumTpl; ++t) {
    p->grid = (uint64_t)ctx->num_cus * (uint64_t)blocks_per_cu(ctx, kind, t);
    p->flat = flat_of(tile_k(kTplChoices[t]));
    tpl = t;
    const uint64_t waves = p->grid * (uint64_t)ctx->wg_waves;
    const uint64_t need =
        ctx->min_tiles_per_wave > 0 ? (uint64_t)ctx->min_tiles_per_wave * waves : waves / 8;
    if (ctx->tpl_index >= 0 || p->flat >= need) break;
  }
  p->tpl = tpl;
  p->nwaves = p->grid * (uint64_t)ctx->wg_waves;
  const uint64_t nwaves = p->nwaves, flat = p->flat;
  p->static_tiles = 0;
  p->nphases = 0;
  if (flat <= 64 * nwaves) {
    p->static_tiles = (flat + nwaves - 1) / nwaves;
  } else {
    const uint64_t g = (uint64_t)ctx->chunks_per_wave;
    uint64_t start = 0, first = 0, left = flat;
    while (g > 0 && p->nphases + 1 < kMaxPhases) {
      const uint64_t chunk = left / (2 * nwaves * g);
      if (chunk < 2 * (uint64_t)ctx->min_chunk || chunk > 0xffffffffull) break;
      p->phases[p->nphases++] = {start, (uint32_t)first, (uint32_t)chunk};
      start += nwaves * g * chunk;
      first += nwaves * g;
      left -= nwaves * g * chunk;
    }
    // the last phase, to the end of the tile space: chunks of min_chunk (fixed chunks: at most
    // min_chunk and at least 8 per wave), or larger where more than 2^30 of them would be needed
    // (chunk indices stay below 2^31 with the waves' last grabs)
    uint64_t cmin = (uint64_t)ctx->min_chunk;
    if (g == 0) {
      const uint64_t per = flat / (8 * nwaves);
      if (per < cmin) cmin = per < 1 ? 1 : per;
    }
    if (left / cmin > (1ull << 30)) cmin = left / (1ull << 30) + 1;
    p->phases[p->nphases++] = {start, (uint32_t)first, (uint32_t)cmin};
  }
  p->replicas = 1;
  while (p->replicas < (uint32_t)ctx->max_replicas &&
         (uint64_t)p->replicas * 2 * nslots * 16 <= nwaves)
    p->replicas *= 2;
}

// The plan's schedule in a call's header.
void plan_to_header(const LaunchPlan& p, LaunchDev* h) {
  h->static_tiles = p.static_tiles;
  h->waves_per_group = (uint32_t)p.wg_waves;
  h->nphases = p.nphases;
  for (uint32_t i = 0; i < p.nphases; ++i) h->phases[i] = p.phases[i];
  h->replicas = p.replicas;
}

// Device-side copies a call leaves behind (clb_normal keeps the reference's Allloop state).
struct Mirrors {
  uint64_t* total = nullptr;  // receives group 0's result
  uint64_t* slot = nullptr;   // receives slot `index`'s shifted value
  uint32_t index = 0;
};

// Reference replay of the BatchCalc jobs `in` (KernelFuncs.as:25-66 with ctx->replay_gws
// work-items per launch): the launches of the reference's double-double kernels (GWS * 16 terms
// each below loopmax32, GWS * 32 above for precision 1), precision 2's range past loopmax32 as an
// exact job (its Refresh kernel is exact per term, so the chunking does not matter), and the tail
// kernel where the range reaches d.  Replay tiles are 64 work-items; *rtiles receives their count.
int expand_replay(const clb_ctx* ctx, const std::vector<HostJob>& in, std::vector<HostJob>* exact,
                  std::vector<ReplayJobDev>* rjobs, std::vector<TailDev>* rtails,
                  uint64_t* rtiles) {
  const uint64_t g = ctx->replay_gws;
  if (ctx->precision != CLB_PREC_U192 && ctx->precision != CLB_PREC_U192_BETA)
    return fail(CLB_EINVAL, "replay: the 106-bit mode (precision 0) is not replayed");
  uint64_t flat = 0;
  for (const HostJob& h : in) {
    if (h.d > (INT64_MAX - 9) / 10)  // the -D OVERDOUBLE kernels need moduli below 2^63
      return fail(CLB_ERANGE, "replay: d = %llu puts moduli at 2^63 or above",
                  (unsigned long long)h.d);
    uint64_t offset = h.k0, k_max = h.k0;
    auto launch = [&](uint64_t end) {
      ReplayJobDev r{};
      r.d = h.d;
      r.offset = offset;
      r.k_max = end;
      r.gsize = g;
      r.flat_begin = flat;
      r.den0 = h.den0;
      r.den1 = h.den1;
      r.numesign = h.numesign;
      r.slot = h.slot;
      const uint64_t items = end - offset < g ? end - offset : g;
      flat += (items + kWave - 1) / kWave;
      rjobs->push_back(r);
    };
    if (h.k0 < h.k1 && h.single) {  // one kernel launch, the precision's kernel for m(k0)
      if (ctx->precision == CLB_PREC_U192_BETA && h.den0 * h.k0 + h.den1 >= (1ull << 32)) {
        HostJob e = h;
        e.tail = false;
        exact->push_back(e);
      } else {
        launch(h.k1);
      }
    } else if (h.k0 < h.k1) {
      const uint64_t x = ((1ull << 32) - h.den1 - 1) / h.den0 + 1;  // KernelFuncs.as:31-33
      const uint64_t lm32 = x < h.k1 ? x : h.k1;
      while (offset < lm32) {  // :34-41
        k_max = k_max + g * 16 < lm32 ? k_max + g * 16 : lm32;
        launch(k_max);
        offset = k_max;
      }
      if (ctx->precision == CLB_PREC_U192_BETA) {
        if (offset < h.k1) {
          HostJob e = h;
          e.k0 = offset;
          e.tail = false;
          exact->push_back(e);
        }
      } else {
        while (offset < h.k1) {  // :45-53
          k_max = k_max + g * 32 < h.k1 ? k_max + g * 32 : h.k1;
          launch(k_max);
          offset = k_max;
        }
      }
    }
    if (h.tail) rtails->push_back({h.d, h.den0, h.den1, h.numesign, h.slot});  // :56-59
  }
  *rtiles = flat;
  return CLB_OK;
}

constexpr size_t kBigSumBytes = (size_t)kRefGroups * 2 * sizeof(double);

// The context's bigSumMem for the 106-bit-mode replay, allocated and cleared on first use.
int r106_state(clb_ctx* ctx, double** out) {
  if (!ctx->r106_state.p) {
    if (int rc = ctx->r106_state.reserve(kBigSumBytes)) return rc;
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(ctx->r106_state.p),
                        kBigSumBytes / 8, ctx->stream));
  }
  *out = reinterpret_cast<double*>(ctx->r106_state.p);
  return CLB_OK;
}

// 3-word result from a double-double: the bit patterns of (hi, lo) and a zero word.
void dd_words(const double* dd, uint64_t w[3]) {
  memcpy(&w[0], &dd[0], 8);
  memcpy(&w[1], &dd[1], 8);
  w[2] = 0;
}

// Reference replay of the 106-bit mode (precision 0): each BatchCalc job on its own bigSumMem
// (cleared, as Set0memdd before every BatchCalc of Allloop and work mode, Bellard_192.hsp:270,
// :311), or, with accumulate_dev (the operators), the one job on the context's bigSumMem with
// bigSum[0] copied to accumulate_dev; then AddAns per group (Sum7 in formula order).  Results are
// double-doubles, in the 3-word outputs as dd_words.  Operator jobs: BatchCalc (clb_batch_calc:
// its launches, the tail where the range reaches d, Ssum), one Sglobal32mtg / Sglobal64mtg launch
// (clb_sglobal: single), Sglobal_after (clb_sglobal_after: single with tail and no head terms),
// Ssum alone (clb_ssum: ssum_op).
int run_replay106(clb_ctx* ctx, const std::vector<HostJob>& hjobs,
                  const std::vector<uint8_t>& shift, const std::vector<GroupDev>& groups,
                  uint64_t* out_host, uint64_t* accumulate_dev, uint64_t* slot_host,
                  const Mirrors& mirrors) {
  const auto t_call = std::chrono::steady_clock::now();
  int rc = set_device(ctx);
  if (rc) return rc;
  const uint64_t gws = ctx->replay_gws;
  const bool op_level = accumulate_dev != nullptr;
  const size_t nslots = hjobs.size();
  uint64_t max_m = 1;
  for (const HostJob& h : hjobs) {
    if (h.d > (INT64_MAX - 9) / 10)
      return fail(CLB_ERANGE, "replay: d = %llu puts moduli at 2^63 or above",
                  (unsigned long long)h.d);
    const uint64_t x32 = ((1ull << 32) - h.den1 - 1) / h.den0 + 1;  // KernelFuncs.as:31-33
    if (h.single && h.k0 < x32 && x32 < h.k1)
      return fail(CLB_EINVAL, "replay: one Sglobal32mtg launch over [%llu, %llu) passes "
                  "loopmax32 = %llu (the reference's 32-bit kernel truncates larger moduli)",
                  (unsigned long long)h.k0, (unsigned long long)h.k1, (unsigned long long)x32);
    if (h.k0 < h.k1 && h.den0 * (h.k1 - 1) + h.den1 > max_m) max_m = h.den0 * (h.k1 - 1) + h.den1;
  }
  if (op_level && nslots != 1) return fail(CLB_EINVAL, "replay: one job per operator call");
  int kind = kPlainW3;
  if ((rc = choose_kind(ctx->kernel, max_m, &kind))) return rc;
  const size_t ngr = groups.size();
  auto align64 = [](size_t x) { return (x + 63) & ~(size_t)63; };
  const size_t off_jobs = 0;
  const size_t off_parts = align64(nslots * sizeof(R106JobDev));
  const size_t off_first = off_parts + align64(nslots * 16);
  const size_t off_count = off_first + align64(ngr * 4);
  const size_t off_shift = off_count + align64(ngr * 4);
  const size_t off_ans = off_shift + align64(nslots);
  const size_t off_slot = off_ans + align64(ngr * 16);
  const size_t bytes = off_slot + nslots * 16;
  if ((rc = ctx->r106_block.reserve(bytes)) || (rc = ctx->r106_host.reserve(bytes))) return rc;
  constexpr size_t kChunk = 64;  // jobs whose 1 MB bigSumMem are live at once
  double* state = nullptr;
  if (op_level && (rc = r106_state(ctx, &state))) return rc;
  if (!op_level && (rc = ctx->r106_bigsum.reserve(kChunk * kBigSumBytes))) return rc;
  uint8_t* dev = ctx->r106_block.p;
  uint8_t* stage = ctx->r106_host.p;
  std::vector<R106JobDev> jobs(nslots);
  for (size_t i = 0; i < nslots; ++i) {
    const HostJob& h = hjobs[i];
    R106JobDev& j = jobs[i];
    j.d = h.d;
    j.k0 = h.k0;
    j.k1 = h.k1;
    j.den0 = h.den0;
    j.den1 = h.den1;
    j.bigsum = (uint64_t)(uintptr_t)(op_level ? reinterpret_cast<uint8_t*>(state)
                                              : ctx->r106_bigsum.p + (i % kChunk) * kBigSumBytes);
    j.part = (uint64_t)(uintptr_t)(dev + off_parts + 16 * i);
    j.numesign = h.numesign;
    j.tail = h.tail;
    j.single = h.single;
    j.ssum = !h.single || h.ssum_op;
  }
  memcpy(stage + off_jobs, jobs.data(), nslots * sizeof(R106JobDev));
  for (size_t g = 0; g < ngr; ++g) {
    memcpy(stage + off_first + 4 * g, &groups[g].first, 4);
    memcpy(stage + off_count + 4 * g, &groups[g].count, 4);
  }
  if (nslots) memcpy(stage + off_shift, shift.data(), nslots);
  memset(stage + off_ans, 0, ngr * 16);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dev, stage, off_slot, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ctx->ev[0], s));
  const R106JobDev* djobs = reinterpret_cast<const R106JobDev*>(dev + off_jobs);
  for (size_t c = 0; c < nslots; c += kChunk) {
    const size_t n = nslots - c < kChunk ? nslots - c : kChunk;
    if (!op_level)
      HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(ctx->r106_bigsum.p),
                          n * kBigSumBytes / 8, s));
    uint64_t most = 0;  // work-groups of the chunk's largest launch
    for (size_t i = c; i < c + n; ++i) {
      const HostJob& h = hjobs[i];
      const uint64_t span = h.k1 > h.k0 ? h.k1 - h.k0 : 0;
      const uint64_t items = span < gws ? span : gws;
      // an empty clb_sglobal still runs group 0 (its sumfix2); Sglobal_after and Ssum run none
      const uint64_t groups = h.ssum_op || (h.single && h.tail) ? 0
                              : h.single && span == 0      ? 1
                                                           : (items + 255) / 256;
      if (groups > most) most = groups;
    }
    HIP_TRY(launch_replay106(kind, djobs + c, (uint32_t)n, gws, (uint32_t)most, s));
  }
  HIP_TRY(hipEventRecord(ctx->ev[1], s));
  const double* dparts = reinterpret_cast<const double*>(dev + off_parts);
  double* dans = reinterpret_cast<double*>(dev + off_ans);
  double* dslot = reinterpret_cast<double*>(dev + off_slot);
  if (op_level) {
    HIP_TRY(hipMemcpyAsync(accumulate_dev, dparts, 16, hipMemcpyDeviceToDevice, s));
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(accumulate_dev + 2), 1, s));
  } else {
    HIP_TRY(launch_replay106_sum7(dparts, reinterpret_cast<const uint32_t*>(dev + off_first),
                                  reinterpret_cast<const uint32_t*>(dev + off_count),
                                  dev + off_shift, dans, dslot, (uint32_t)ngr, s));
    if (mirrors.total) {
      HIP_TRY(hipMemcpyAsync(mirrors.total, dans, 16, hipMemcpyDeviceToDevice, s));
      HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(mirrors.total + 2), 1, s));
    }
    if (mirrors.slot) {
      HIP_TRY(hipMemcpyAsync(mirrors.slot, dslot + 2 * mirrors.index, 16,
                             hipMemcpyDeviceToDevice, s));
      HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(mirrors.slot + 2), 1, s));
    }
    HIP_TRY(hipMemcpyAsync(stage + off_ans, dev + off_ans, bytes - off_ans,
                           hipMemcpyDeviceToHost, s));
  }
  hipError_t e2 = hipStreamSynchronize(s);
  float head = 0.f;
  if (e2 == hipSuccess) e2 = hipEventElapsedTime(&head, ctx->ev[0], ctx->ev[1]);
  if (e2 != hipSuccess) return fail(CLB_EDEVICE, "replay failed: %s", hipGetErrorString(e2));
  if (!op_level) {
    const double* ans = reinterpret_cast<const double*>(stage + off_ans);
    const double* sv = reinterpret_cast<const double*>(stage + off_slot);
    if (out_host)
      for (size_t g = 0; g < ngr; ++g) dd_words(ans + 2 * g, out_host + 3 * g);
    if (slot_host)
      for (size_t i = 0; i < nslots; ++i) dd_words(sv + 2 * i, slot_host + 3 * i);
  }
  ctx->head_ms = head;
  ctx->clk_valid = false;
  ctx->total_ms = (float)std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t_call)
                      .count();
  return CLB_OK;
}

// Runs head jobs + tails into slots [0, nslots), then combines the slots by groups.
// out_host receives 3 words per group unless accumulate_dev is given, in which case the single
// group's value is added into that device buffer.  slot_host (optional) receives each slot's
// shifted value.
int run(clb_ctx* ctx, const std::vector<HostJob>& hjobs, const std::vector<uint8_t>& shift,
        const std::vector<GroupDev>& groups, uint64_t* out_host, uint64_t* accumulate_dev,
        uint64_t* slot_host, const Mirrors& mirrors = Mirrors()) {
  if (ctx->replay_gws && ctx->precision == CLB_PREC_DD106)
    return run_replay106(ctx, hjobs, shift, groups, out_host, accumulate_dev, slot_host, mirrors);
  const auto t_call = std::chrono::steady_clock::now();
  int rc = set_device(ctx);
  if (rc) return rc;
  const size_t nslots = shift.size();
  std::vector<JobDev> jobs;
  std::vector<TailDev> tails;
  // Reference replay: the head jobs and tails below are then only the exact part.
  std::vector<HostJob> exact_jobs;
  std::vector<ReplayJobDev> rjobs;
  std::vector<TailDev> rtails;
  uint64_t rtiles = 0;
  const bool replay = ctx->replay_gws != 0;
  if (replay && (rc = expand_replay(ctx, hjobs, &exact_jobs, &rjobs, &rtails, &rtiles))) return rc;
  const std::vector<HostJob>& head_jobs = replay ? exact_jobs : hjobs;
  uint64_t max_m = 1;
  size_t nhead = 0;
  for (const HostJob& h : head_jobs) {
    // exponents 10 (d - k) must fit 64 bits; tail moduli den0 (d + 19) + den1 below 2^63 (the
    // bit-serial division's bound)
    if (h.d > UINT64_MAX / 10) return fail(CLB_ERANGE, "loop count %llu too large",
                                           (unsigned long long)h.d);
    if (h.tail) {
      if (h.d + kTailTerms > ((1ull << 63) - h.den1) / h.den0)
        return fail(CLB_ERANGE, "tail modulus of d = %llu exceeds 2^63", (unsigned long long)h.d);
      tails.push_back({h.d, h.den0, h.den1, h.numesign, h.slot});
    }
    if (h.k0 >= h.k1) continue;
    const uint64_t m = h.den0 * (h.k1 - 1) + h.den1;
    if (m > max_m) max_m = m;
    ++nhead;
  }
  int kind = kSplitW3;
  if (nhead && (rc = choose_kind(ctx->kernel, max_m, &kind))) return rc;
  if (nhead > 0x7fffffff || tails.size() > 0x7fffffff || groups.size() > 0x7fffffff)
    return fail(CLB_EINVAL, "too many jobs");
  LaunchPlan plan;
  plan_launch(ctx, kind, nslots, [&](uint64_t tk) {
    uint64_t f = 0;
    for (const HostJob& h : head_jobs)
      if (h.k0 < h.k1) f += (h.k1 + tk - 1) / tk - h.k0 / tk;
    return f;
  }, &plan);
  const int tpl = plan.tpl;
  const uint64_t grid = plan.grid;
  uint64_t flat = 0;
  const uint64_t tk = tile_k(kTplChoices[tpl]);
  for (const HostJob& h : head_jobs) {
    if (h.k0 >= h.k1) continue;
    JobDev j{};
    j.d = h.d;
    j.k0 = h.k0;
    j.k1 = h.k1;
    j.tile_first = h.k0 / tk;
    j.flat_begin = flat;
    j.den0 = h.den0;
    j.den1 = h.den1;
    j.numesign = h.numesign;
    j.slot = h.slot;
    flat += (h.k1 + tk - 1) / tk - j.tile_first;
    jobs.push_back(j);
  }
  const uint32_t replicas = plan.replicas;
  // The block: header, jobs, tails, groups, shifts, accumulator (copied in as zeros), results
  // (group results, then slot values), each 64-byte aligned.
  auto align64 = [](size_t x) { return (x + 63) & ~(size_t)63; };
  const size_t off_jobs = align64(sizeof(LaunchDev));
  const size_t off_tails = off_jobs + align64(jobs.size() * sizeof(JobDev));
  const size_t off_rjobs = off_tails + align64(tails.size() * sizeof(TailDev));
  const size_t off_rtails = off_rjobs + align64(rjobs.size() * sizeof(ReplayJobDev));
  const size_t off_groups = off_rtails + align64(rtails.size() * sizeof(TailDev));
  const size_t off_shifts = off_groups + align64(groups.size() * sizeof(GroupDev));
  const size_t off_acc = off_shifts + align64(nslots);
  const size_t off_out = off_acc + nslots * replicas * kAccWords * sizeof(uint64_t);
  const size_t out_words = 3 * groups.size() + (slot_host ? 3 * nslots : 0);
  const size_t bytes = off_out + out_words * sizeof(uint64_t);
  if (bytes > 0xffffffffull) return fail(CLB_EINVAL, "too many jobs for one call");
  if ((rc = ctx->block.reserve(bytes)) || (rc = ctx->block_host.reserve(bytes))) return rc;
  uint8_t* dev = ctx->block.p;
  LaunchDev h{};
  h.counter = 0;
  h.total_tiles = flat;
  plan_to_header(plan, &h);
  h.dst = (uint64_t)(uintptr_t)(accumulate_dev ? accumulate_dev
                                               : reinterpret_cast<uint64_t*>(dev + off_out));
  h.slot_out = slot_host ? (uint64_t)(uintptr_t)(dev + off_out + 3 * groups.size() * 8) : 0;
  h.mirror_total = (uint64_t)(uintptr_t)mirrors.total;
  h.mirror_slot = (uint64_t)(uintptr_t)mirrors.slot;
  h.mirror_index = mirrors.index;
  h.clk_out = (uint64_t)(uintptr_t)ctx->clk;
  h.accumulate = accumulate_dev != nullptr;
  h.njobs = (uint32_t)jobs.size();
  h.ntails = (uint32_t)tails.size();
  h.ngroups = (uint32_t)groups.size();
  h.nslots = (uint32_t)nslots;
  h.off_jobs = (uint32_t)off_jobs;
  h.off_tails = (uint32_t)off_tails;
  h.off_groups = (uint32_t)off_groups;
  h.off_shifts = (uint32_t)off_shifts;
  h.off_acc = (uint32_t)off_acc;
  h.off_out = (uint32_t)off_out;
  int rgrid = 0, rkind = kPlainW3;
  if (replay) {
    uint64_t rmax_m = 1;
    for (const ReplayJobDev& r : rjobs) {
      const uint64_t m = r.den0 * (r.k_max - 1) + r.den1;
      if (m > rmax_m) rmax_m = m;
    }
    if ((rc = choose_kind(ctx->kernel, rmax_m, &rkind))) return rc;
    int& res = ctx->replay_resident[rkind];
    if (res < 0) res = replay_blocks_per_cu(rkind);
    if (res <= 0) return fail(CLB_EDEVICE, "replay kernel: no occupancy");
    rgrid = ctx->num_cus * res;
    const uint64_t rwaves = (uint64_t)rgrid * (kBlock / kWave);
    const uint64_t per = rtiles / (8 * rwaves);
    if (rjobs.size() > 0x7fffffff || rtails.size() > 0x7fffffff)
      return fail(CLB_EINVAL, "too many replayed launches");
    h.rcounter = 0;
    h.rtotal_tiles = rtiles;
    h.nrjobs = (uint32_t)rjobs.size();
    h.nrtails = (uint32_t)rtails.size();
    h.rchunk = (uint32_t)(per < 1 ? 1 : per > 256 ? 256 : per);
  }
  h.off_rjobs = (uint32_t)off_rjobs;
  h.off_rtails = (uint32_t)off_rtails;
  uint8_t* stage = ctx->block_host.p;
  memcpy(stage, &h, sizeof h);
  if (!jobs.empty()) memcpy(stage + off_jobs, jobs.data(), jobs.size() * sizeof(JobDev));
  if (!tails.empty()) memcpy(stage + off_tails, tails.data(), tails.size() * sizeof(TailDev));
  if (!rjobs.empty())
    memcpy(stage + off_rjobs, rjobs.data(), rjobs.size() * sizeof(ReplayJobDev));
  if (!rtails.empty())
    memcpy(stage + off_rtails, rtails.data(), rtails.size() * sizeof(TailDev));
  if (!groups.empty()) memcpy(stage + off_groups, groups.data(), groups.size() * sizeof(GroupDev));
  if (nslots) memcpy(stage + off_shifts, shift.data(), nslots);
  memset(stage + off_acc, 0, off_out - off_acc);
  LaunchDev* hdr = reinterpret_cast<LaunchDev*>(dev);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dev, stage, off_out, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ctx->ev[0], s));
  if (ctx->use_graph && !replay) {
    hipGraphExec_t exec = nullptr;
    if ((rc = graph_for(ctx, kind, tpl, (int)grid, &exec))) return rc;
    HIP_TRY(hipGraphLaunch(exec, s));
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
  } else {
    if (!replay || !jobs.empty())
      HIP_TRY(launch_head(kind, tpl, hdr, (int)grid, ctx->wg_waves, s));
    if (replay && (!rjobs.empty() || !rtails.empty()))
      HIP_TRY(launch_replay(rkind, hdr, rgrid, s));
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
    HIP_TRY(launch_combine(hdr, s));
  }
  const size_t back = accumulate_dev ? (slot_host ? 3 * nslots : 0) : out_words;
  const size_t back_off = accumulate_dev ? off_out + 3 * groups.size() * 8 : off_out;
  if (back) HIP_TRY(hipMemcpyAsync(stage + back_off, dev + back_off, back * 8,
                                   hipMemcpyDeviceToHost, s));
  hipError_t e2 = hipStreamSynchronize(s);
  float head = 0.f;
  if (e2 == hipSuccess) e2 = hipEventElapsedTime(&head, ctx->ev[0], ctx->ev[1]);
  if (e2 != hipSuccess)
    return fail(CLB_EDEVICE, "term launch failed: %s", hipGetErrorString(e2));
  const uint64_t* res = reinterpret_cast<const uint64_t*>(stage + off_out);
  if (out_host && !accumulate_dev && !groups.empty())
    memcpy(out_host, res, 3 * groups.size() * sizeof(uint64_t));
  if (slot_host && nslots)
    memcpy(slot_host, res + 3 * groups.size(), 3 * nslots * sizeof(uint64_t));
  ctx->head_ms = head;
  ctx->clk_valid = !replay || !jobs.empty();  // k_head ran
  ctx->total_ms = (float)std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t_call)
                      .count();
  return CLB_OK;
}

// KernelFuncs.as:25-66 semantics: head = [k_start, min(k_end, d)), tail iff the final offset
// (the end of the head if it ran, else k_start) equals d.
HostJob batch_job(uint64_t d, uint64_t k_start, uint64_t k_end, uint64_t den0, uint64_t den1,
                  uint32_t numesign, uint32_t slot) {
  const uint64_t hi = k_end < d ? k_end : d;
  HostJob h{};
  h.d = d;
  h.den0 = den0;
  h.den1 = den1;
  h.numesign = numesign;
  h.slot = slot;
  uint64_t offset = k_start;
  if (k_start < hi) {
    h.k0 = k_start;
    h.k1 = hi;
    offset = hi;
  } else {
    h.k0 = h.k1 = 0;
  }
  h.tail = (offset == d);
  return h;
}

int check_ctx(clb_ctx* ctx) {
  if (!ctx) return fail(CLB_EINVAL, "null context");
  return CLB_OK;
}

// How a result computed in `precision` on ctx is printed: the 106-bit-mode replay returns
// double-double bits.
int out_format(const clb_ctx* ctx, int precision) {
  return ctx->replay_gws && precision == CLB_PREC_DD106 ? 3 : precision;
}

// ------------------------------------------------------------------ work mode helpers

// Sum of bitlen(10 x) over x in [a, b] (a >= 1): the exponent lengths 10 (d - k) of the head
// terms k in [d - b, d - a], i.e. the modular squares the term kernels spend on them.
double sum_bitlen_10x(uint64_t a, uint64_t b) {
  if (a > b) return 0.0;
  double total = 0.0;
  for (int bl = 1; bl <= 68; ++bl) {
    const unsigned __int128 lo_v = ((unsigned __int128)1 << (bl - 1)) + 9;  // ceil(2^(bl-1) / 10)
    const unsigned __int128 hi_v = ((unsigned __int128)1 << bl) - 1;
    const unsigned __int128 lo = lo_v / 10 > a ? lo_v / 10 : a;
    const unsigned __int128 hi = hi_v / 10 < b ? hi_v / 10 : b;
    if (hi >= lo) total += (double)bl * (double)(uint64_t)(hi - lo + 1);
  }
  return total;
}

// Cost of one BatchCalc job in modular squares: bitlen(10 (d - k)) squares plus about three
// squares' worth of quotient and bookkeeping per head term, and the tail's twenty terms.
double job_squares(uint64_t d, uint64_t k_begin, uint64_t k_end) {
  const uint64_t hi = k_end < d ? k_end : d;
  double c = 60.0;
  if (k_begin < hi) c += sum_bitlen_10x(d - hi + 1, d - k_begin) + 3.0 * (double)(hi - k_begin);
  return c;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int clb_get_dev_count(int* count) {
  if (!count) return fail(CLB_EINVAL, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(CLB_EDEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *count = n;
  return CLB_OK;
}

int clb_get_device_name(int device, char* name, size_t cap) {
  if (!name || cap == 0) return fail(CLB_EINVAL, "null name buffer");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  snprintf(name, cap, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
  return CLB_OK;
}

int clb_ctx_create(int device, int precision, clb_ctx** out) {
  if (!out) return fail(CLB_EINVAL, "null output");
  *out = nullptr;
  if (precision < CLB_PREC_DD106 || precision > CLB_PREC_U192_BETA)
    return fail(CLB_EINVAL, "precision must be 0, 1 or 2");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(CLB_EDEVICE, "no HIP device available");  // Bellard_192.hsp:17
  if (device < 0 || device >= n) return fail(CLB_EINVAL, "device %d of %d", device, n);
  clb_ctx* ctx = new clb_ctx();
  ctx->device = device;
  ctx->precision = precision;
  auto cleanup = [&](int rc) {
    clb_ctx_destroy(ctx);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess)
    return cleanup(fail(CLB_EDEVICE, "device %d: %s", device, hipGetErrorString(e)));
  ctx->num_cus = prop.multiProcessorCount;
  snprintf(ctx->device_name, sizeof ctx->device_name, "%s",
           prop.name[0] ? prop.name : prop.gcnArchName);
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
    return cleanup(fail(CLB_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e)));
  // The no-allocation entry point's device array and staging, once per device (the device is
  // current); a failure only disables that entry point, which then reports it.
  ctx->noalloc_err = noalloc_setup(device);
  for (auto& ev : ctx->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess)
      return cleanup(fail(CLB_EDEVICE, "hipEventCreate: %s", hipGetErrorString(e)));
  if ((e = hipMalloc((void**)&ctx->partial, 3 * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc((void**)&ctx->ans, 3 * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc((void**)&ctx->clk, 4 * kClockXcds * sizeof(uint64_t))) != hipSuccess)
    return cleanup(fail(CLB_ENOMEM, "hipMalloc: %s", hipGetErrorString(e)));
  // The doubling squares (term kernels and replay kernels, two code objects) need the FP64
  // output modifier, which the hardware applies only in device code built with the IEEE mode bit
  // clear (clbellard_amd/build.py): refuse a build whose kernels would silently compute other
  // numbers.
  uint64_t probe[4] = {0, 0, 0, 0};
  if ((e = launch_selfcheck(ctx->partial, ctx->stream)) != hipSuccess ||
      (e = launch_replay_selfcheck(ctx->ans, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(probe, ctx->partial, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(probe + 2, ctx->ans, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return cleanup(fail(CLB_EDEVICE, "self-check: %s", hipGetErrorString(e)));
  for (int i = 0; i < 4; i += 2)
    if (probe[i] != 18 || probe[i + 1] != 9)
      return cleanup(fail(CLB_EDEVICE,
                          "self-check: the %s device code does not apply FP64 output modifiers "
                          "(got %llu, %llu for 18, 9); build it with -mno-amdgpu-ieee",
                          i == 0 ? "term-kernel" : "replay", (unsigned long long)probe[i],
                          (unsigned long long)probe[i + 1]));
  if ((e = hipMemsetAsync(ctx->partial, 0, 3 * sizeof(uint64_t), ctx->stream)) != hipSuccess ||
      (e = hipMemsetAsync(ctx->ans, 0, 3 * sizeof(uint64_t), ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return cleanup(fail(CLB_EDEVICE, "clear: %s", hipGetErrorString(e)));
  *out = ctx;
  return CLB_OK;
}

int clb_ctx_destroy(clb_ctx* ctx) {
  if (!ctx) return CLB_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->block.release();
  ctx->block_host.release();
  ctx->r106_block.release();
  ctx->r106_host.release();
  ctx->r106_bigsum.release();
  ctx->r106_state.release();
  for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
  if (ctx->partial) (void)hipFree(ctx->partial);
  if (ctx->ans) (void)hipFree(ctx->ans);
  if (ctx->clk) (void)hipFree(ctx->clk);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CLB_OK;
}

int clb_set_kernel(clb_ctx* ctx, int kernel) {
  if (int rc = check_ctx(ctx)) return rc;
  if (kernel < CLB_KERNEL_AUTO || kernel > CLB_KERNEL_SPLIT)
    return fail(CLB_EINVAL, "kernel must be 0 (auto), 1 (plain) or 2 (split)");
  ctx->kernel = kernel;
  return CLB_OK;
}

int clb_set_tiling(clb_ctx* ctx, int terms_per_lane, int min_tiles_per_wave) {
  if (int rc = check_ctx(ctx)) return rc;
  int idx = -1;
  for (int t = 0; t < kNumTpl; ++t)
    if (kTplChoices[t] == terms_per_lane) idx = t;
  if (terms_per_lane != 0 && idx < 0)
    return fail(CLB_EINVAL, "terms per lane must be 0, 16, 4 or 1");
  if (min_tiles_per_wave < 0 || min_tiles_per_wave > 1024)
    return fail(CLB_EINVAL, "min_tiles_per_wave 0..1024");
  ctx->tpl_index = idx;
  ctx->min_tiles_per_wave = min_tiles_per_wave;
  return CLB_OK;
}

int clb_autotune(clb_ctx* ctx, uint64_t n, double budget_ms, int* terms_per_lane,
                 int* blocks_per_cu_out) {
  if (int rc = check_ctx(ctx)) return rc;
  uint64_t d;
  if (int rc = normalize(n, nullptr, &d, nullptr)) return rc;
  if (!(budget_ms > 0.0)) return fail(CLB_EINVAL, "budget_ms must be positive");
  const int saved_tpl = ctx->tpl_index, saved_bpc = ctx->blocks_per_cu;
  const int saved_min = ctx->min_tiles_per_wave;
  const int saved_wg = ctx->wg_waves, saved_chunk = ctx->min_chunk;
  const int saved_cpw = ctx->chunks_per_wave, saved_rep = ctx->max_replicas;
  const bool saved_graph = ctx->use_graph;
  const uint64_t saved_replay = ctx->replay_gws;
  ctx->use_graph = false;
  ctx->replay_gws = 0;  // the knobs belong to the exact kernels, so those are timed
  // The sample: the seven formulas' head terms from k0, s terms each (all of [0, d) when that
  // fits the budget), no tails.
  auto jobs_for = [&](uint64_t k0, uint64_t s) {
    std::vector<HostJob> jobs;
    for (int j = 0; j < 7; ++j) {
      HostJob h =
          batch_job(d, k0, k0 + s, kDen0[j], kDen1[j], (uint32_t)kNumeSign[j], (uint32_t)j);
      h.tail = false;
      jobs.push_back(h);
    }
    return jobs;
  };
  std::vector<uint8_t> shift(kShift, kShift + 7);
  std::vector<GroupDev> groups{{0, 7}};
  uint64_t res[3], best_res[3] = {0, 0, 0};
  auto restore = [&](int rc) {
    ctx->tpl_index = saved_tpl;
    ctx->blocks_per_cu = saved_bpc;
    ctx->min_tiles_per_wave = saved_min;
    ctx->wg_waves = saved_wg;
    ctx->min_chunk = saved_chunk;
    ctx->chunks_per_wave = saved_cpw;
    ctx->max_replicas = saved_rep;
    ctx->use_graph = saved_graph;
    ctx->replay_gws = saved_replay;
    return rc;
  };
  // probe at the default settings to size the sample
  ctx->tpl_index = 0;
  ctx->blocks_per_cu = 0;
  ctx->wg_waves = 4;
  ctx->min_chunk = 256;
  ctx->chunks_per_wave = 0;
  ctx->max_replicas = 64;
  const uint64_t probe = d < (1ull << 20) ? d : (1ull << 20);
  const uint64_t k_mid = d > probe ? (d - probe) / 2 : 0;
  if (int rc = run(ctx, jobs_for(k_mid, probe), shift, groups, res, nullptr, nullptr))
    return restore(rc);
  const double per_term = (ctx->head_ms > 0.f ? ctx->head_ms : 1e-3) / (double)probe;
  const int resident = blocks_per_cu(ctx, kSplitW3, 0);
  const int bpc_choices[3] = {resident, (3 * resident + 3) / 4, (resident + 1) / 2};
  // Stage 2 changes one knob at a time on the best settings so far (a value that won stays while
  // the next one is tried): the workgroup size (1, 2, 8 waves), the tile queue (fixed chunks of 64
  // or 1024 tiles, a guided queue) and the slot copies (at most 4).
  struct Cfg {
    int tpl, bpc, wg, chunk, cpw, rep;  // bpc: resident workgroups of 4 waves per CU
  };
  enum Knob { kWorkgroup, kQueue, kCopies };
  struct Step {
    Knob knob;
    int a, b;  // workgroup waves; chunk tiles and chunks per wave (guided); slot copies
  };
  const Step stage2[] = {{kWorkgroup, 1, 0}, {kWorkgroup, 2, 0}, {kWorkgroup, 8, 0},
                         {kQueue, 64, 0},    {kQueue, 1024, 0},  {kQueue, 64, 4},
                         {kCopies, 4, 0}};
  const int nstage2 = (int)(sizeof stage2 / sizeof stage2[0]);
  // Each trial times the candidate and the best so far alternately, reps times each: the GPU's
  // clock moves over a tuning run (it settles under load), and timing candidates one after the
  // other let whichever came later look faster.
  const int ncand = kNumTpl * 3 + nstage2, reps = 3;
  uint64_t s = (uint64_t)(budget_ms / (2 * ncand * reps + 1) / per_term);
  if (s < probe) s = probe;
  if (s > d) s = d;
  const uint64_t k0 = (d - s) / 2;
  const std::vector<HostJob> sample = jobs_for(k0, s);
  auto apply = [&](const Cfg& c) {
    ctx->tpl_index = c.tpl;
    ctx->wg_waves = c.wg;
    // the same resident waves per CU: workgroups of 4 waves per CU rescaled to this size
    ctx->blocks_per_cu = c.bpc == resident ? 0 : (c.bpc * 4 + c.wg - 1) / c.wg;
    ctx->min_chunk = c.chunk;
    ctx->chunks_per_wave = c.cpw;
    ctx->max_replicas = c.rep;
  };
  // the default first: 16 terms per lane, full residency, 256 work-items per workgroup, fixed
  // chunks of 256 tiles, up to 64 slot copies
  Cfg best = {0, resident, 4, 256, 0, 64};
  bool have_res = false;
  // A candidate replaces the best only when it is faster by 2% in the interleaved timing, so
  // timing noise does not move the choice; every launch must return the same bits.
  auto trial = [&](const Cfg& cand, bool* better) -> int {
    double t_best = -1.0, t_cand = -1.0;
    for (int r = 0; r < reps; ++r) {
      for (int which = 0; which < 2; ++which) {
        apply(which ? cand : best);
        if (int rc = run(ctx, sample, shift, groups, res, nullptr, nullptr)) return rc;
        double& t = which ? t_cand : t_best;
        if (t < 0.0 || ctx->head_ms < t) t = ctx->head_ms;
        if (have_res && memcmp(res, best_res, sizeof res) != 0)
          return fail(CLB_EDEVICE, "autotune: results differ between launch shapes");
        memcpy(best_res, res, sizeof res);
        have_res = true;
      }
    }
    *better = t_cand < t_best * 0.98;
    return CLB_OK;
  };
  for (int t = 0; t < kNumTpl; ++t)
    for (int b = 0; b < 3; ++b) {
      if (b > 0 && bpc_choices[b] == bpc_choices[b - 1]) continue;
      if (t == 0 && b == 0) continue;  // the default itself
      const Cfg cand = {t, bpc_choices[b], 4, 256, 0, 64};
      bool better = false;
      if (int rc = trial(cand, &better)) return restore(rc);
      if (better) best = cand;
    }
  for (int i = 0; i < nstage2; ++i) {
    const Step& st = stage2[i];
    Cfg cand = best;
    if (st.knob == kWorkgroup) {
      cand.wg = st.a;
    } else if (st.knob == kQueue) {
      cand.chunk = st.a;
      cand.cpw = st.b;
    } else {
      cand.rep = st.a;
    }
    bool better = false;
    if (int rc = trial(cand, &better)) return restore(rc);
    if (better) best = cand;
  }
  restore(CLB_OK);
  apply(best);
  if (terms_per_lane) *terms_per_lane = kTplChoices[best.tpl];
  if (blocks_per_cu_out) *blocks_per_cu_out = best.bpc;
  return CLB_OK;
}

int clb_get_schedule(clb_ctx* ctx, int* terms_per_lane, int* blocks_per_cu_out,
                     int* waves_per_group, int* min_chunk, int* chunks_per_wave,
                     int* max_replicas) {
  if (int rc = check_ctx(ctx)) return rc;
  if (terms_per_lane) *terms_per_lane = ctx->tpl_index < 0 ? 0 : kTplChoices[ctx->tpl_index];
  if (blocks_per_cu_out) *blocks_per_cu_out = ctx->blocks_per_cu;
  if (waves_per_group) *waves_per_group = ctx->wg_waves;
  if (min_chunk) *min_chunk = ctx->min_chunk;
  if (chunks_per_wave) *chunks_per_wave = ctx->chunks_per_wave;
  if (max_replicas) *max_replicas = ctx->max_replicas;
  return CLB_OK;
}

int clb_save_schedule(clb_ctx* ctx, const char* path) {
  if (int rc = check_ctx(ctx)) return rc;
  clb_schedule s;
  memset(&s, 0, sizeof s);
  snprintf(s.device, sizeof s.device, "%s", ctx->device_name);
  s.terms_per_lane = ctx->tpl_index < 0 ? 0 : kTplChoices[ctx->tpl_index];
  s.min_tiles_per_wave = ctx->min_tiles_per_wave;
  s.blocks_per_cu = ctx->blocks_per_cu;
  s.waves_per_group = ctx->wg_waves;
  s.min_chunk = ctx->min_chunk;
  s.chunks_per_wave = ctx->chunks_per_wave;
  s.max_replicas = ctx->max_replicas;
  return clb_write_schedule(path, &s);
}

int clb_load_schedule(clb_ctx* ctx, const char* path) {
  if (int rc = check_ctx(ctx)) return rc;
  clb_schedule s;
  if (int rc = clb_read_schedule(path, &s)) return rc;
  if (strcmp(s.device, ctx->device_name) != 0)
    return fail(CLB_EINVAL, "%s holds settings for '%s', not for this context's '%s'", path,
                s.device, ctx->device_name);
  // clb_read_schedule checked every value against these setters' ranges: none of them fails
  int rc = clb_set_tiling(ctx, s.terms_per_lane, s.min_tiles_per_wave);
  if (!rc) rc = clb_set_launch(ctx, s.blocks_per_cu);
  if (!rc)
    rc = clb_set_schedule(ctx, s.waves_per_group, s.min_chunk, s.chunks_per_wave, s.max_replicas);
  return rc;
}

int clb_set_replay(clb_ctx* ctx, uint64_t global_work_size) {
  if (int rc = check_ctx(ctx)) return rc;
  if (global_work_size == 0) {
    ctx->replay_gws = 0;
    return CLB_OK;
  }
  if (global_work_size % 256 || global_work_size > kRefGlobalWorkSize)
    return fail(CLB_EINVAL, "replay: %llu work-items (a multiple of 256 up to 2^24)",
                (unsigned long long)global_work_size);
  ctx->replay_gws = global_work_size;
  return CLB_OK;
}

int clb_set_graph(clb_ctx* ctx, int enable) {
  if (int rc = check_ctx(ctx)) return rc;
  if (enable != 0 && enable != 1) return fail(CLB_EINVAL, "graph mode must be 0 or 1");
  ctx->use_graph = enable != 0;
  return CLB_OK;
}

int clb_set_launch(clb_ctx* ctx, int blocks_per_cu) {
  if (int rc = check_ctx(ctx)) return rc;
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(CLB_EINVAL, "blocks_per_cu 0..64");
  ctx->blocks_per_cu = blocks_per_cu;
  return CLB_OK;
}

int clb_set_schedule(clb_ctx* ctx, int waves_per_group, int min_chunk, int chunks_per_wave,
                     int max_replicas) {
  if (int rc = check_ctx(ctx)) return rc;
  if (waves_per_group != 0 && waves_per_group != 1 && waves_per_group != 2 &&
      waves_per_group != 4 && waves_per_group != 8)
    return fail(CLB_EINVAL, "waves_per_group must be 0, 1, 2, 4 or 8");
  if (min_chunk < 0 || min_chunk > 65536) return fail(CLB_EINVAL, "min_chunk 0..65536");
  if (chunks_per_wave < -1 || chunks_per_wave > 64)
    return fail(CLB_EINVAL, "chunks_per_wave -1..64");
  if (max_replicas < 0 || max_replicas > 64 || (max_replicas & (max_replicas - 1)))
    return fail(CLB_EINVAL, "max_replicas 0 or a power of 2 up to 64");
  ctx->wg_waves = waves_per_group ? waves_per_group : 4;
  ctx->min_chunk = min_chunk ? min_chunk : 256;
  ctx->chunks_per_wave = chunks_per_wave < 0 ? 0 : chunks_per_wave;
  ctx->max_replicas = max_replicas ? max_replicas : 64;
  return CLB_OK;
}

int clb_set0mem(clb_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = set_device(ctx)) return rc;
  HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(ctx->partial), 3, ctx->stream));
  if (ctx->r106_state.p)  // the 106-bit-mode replay's bigSumMem (Set0memdd, KernelFuncs.as:20-23)
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(ctx->r106_state.p),
                        kBigSumBytes / 8, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CLB_OK;
}

int clb_clear_ans(clb_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = set_device(ctx)) return rc;
  HIP_TRY(launch_zero(reinterpret_cast<unsigned long long*>(ctx->ans), 3, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CLB_OK;
}

static int run_into_partial(clb_ctx* ctx, const HostJob& h) {
  std::vector<HostJob> jobs{h};
  std::vector<uint8_t> shift{0};
  std::vector<GroupDev> groups{{0, 1}};
  return run(ctx, jobs, shift, groups, nullptr, ctx->partial, nullptr);
}

int clb_batch_calc(clb_ctx* ctx, uint64_t d, uint64_t k_start, uint64_t k_end, int j) {
  if (int rc = check_ctx(ctx)) return rc;
  if (j < 0 || j > 6) return fail(CLB_EINVAL, "formula %d (need 0..6)", j);
  if (d == 0) return fail(CLB_EINVAL, "d must be positive");
  return run_into_partial(ctx, batch_job(d, k_start, k_end, kDen0[j], kDen1[j],
                                         (uint32_t)kNumeSign[j], 0));
}

static int check_formula_args(uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1) {
  if (d == 0) return fail(CLB_EINVAL, "d must be positive");
  if (numesign != 0 && numesign != 1) return fail(CLB_EINVAL, "numesign must be 0 or 1");
  if (den0 == 0 || (den0 & 1) || !(den1 & 1))
    return fail(CLB_EINVAL, "denominator den0*k+den1 must be odd with den0 > 0 even");
  if (d > (UINT64_MAX - den1) / den0 / 2) return fail(CLB_ERANGE, "modulus overflows");
  return CLB_OK;
}

int clb_sglobal(clb_ctx* ctx, uint64_t offset, uint64_t d, uint64_t k_max, int64_t numesign,
                uint64_t den0, uint64_t den1) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = check_formula_args(d, numesign, den0, den1)) return rc;
  if (k_max > d) return fail(CLB_EINVAL, "k_max %llu beyond d %llu", (unsigned long long)k_max,
                             (unsigned long long)d);
  HostJob h{};
  h.d = d;
  h.den0 = den0;
  h.den1 = den1;
  h.numesign = (uint32_t)numesign;
  h.k0 = offset < k_max ? offset : 0;
  h.k1 = offset < k_max ? k_max : 0;
  h.tail = false;
  h.single = true;
  return run_into_partial(ctx, h);
}

int clb_sglobal_after(clb_ctx* ctx, uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1) {
  if (int rc = check_ctx(ctx)) return rc;
  if (int rc = check_formula_args(d, numesign, den0, den1)) return rc;
  HostJob h{};
  h.d = d;
  h.den0 = den0;
  h.den1 = den1;
  h.numesign = (uint32_t)numesign;
  h.tail = true;
  h.single = true;
  return run_into_partial(ctx, h);
}

int clb_ssum(clb_ctx* ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  // Outside the 106-bit-mode replay the operators add straight into the partial (Ssum_192's
  // 192-bit sums are exact, so their order does not matter): nothing to do.
  if (!(ctx->replay_gws && ctx->precision == CLB_PREC_DD106)) return CLB_OK;
  HostJob h{};
  h.d = 1;
  h.den0 = 2;
  h.den1 = 1;
  h.tail = false;
  h.single = true;
  h.ssum_op = true;
  return run_into_partial(ctx, h);
}

int clb_add_ans(clb_ctx* ctx, int j) {
  if (int rc = check_ctx(ctx)) return rc;
  if (j < 0 || j > 6) return fail(CLB_EINVAL, "formula %d (need 0..6)", j);
  if (int rc = set_device(ctx)) return rc;
  if (ctx->replay_gws && ctx->precision == CLB_PREC_DD106) {
    // Sum7 (Bellard.cl:1324-1356) on the context's bigSum[0] and dev_ans, double-doubles.
    double* state = nullptr;
    if (int rc = r106_state(ctx, &state)) return rc;
    if (int rc = ctx->r106_block.reserve(64)) return rc;
    if (int rc = ctx->r106_host.reserve(64)) return rc;
    const uint32_t tab[2] = {0, 1};
    memcpy(ctx->r106_host.p, tab, 8);
    ctx->r106_host.p[8] = kShift[j];
    uint8_t* dev = ctx->r106_block.p;
    HIP_TRY(hipMemcpyAsync(dev, ctx->r106_host.p, 9, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_replay106_sum7(state, reinterpret_cast<const uint32_t*>(dev),
                                  reinterpret_cast<const uint32_t*>(dev + 4), dev + 8,
                                  reinterpret_cast<double*>(ctx->ans), state, 1, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->partial, state, 16, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
  }
  HIP_TRY(launch_sum7(ctx->ans, ctx->partial, kShift[j], ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CLB_OK;
}

static int read3(clb_ctx* ctx, const uint64_t* dev, uint64_t w[3]) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!w) return fail(CLB_EINVAL, "null output");
  if (int rc = set_device(ctx)) return rc;
  HIP_TRY(hipMemcpyAsync(w, dev, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CLB_OK;
}

int clb_read_partial(clb_ctx* ctx, uint64_t w[3]) {
  return read3(ctx, ctx ? ctx->partial : nullptr, w);
}

int clb_read_ans(clb_ctx* ctx, uint64_t w[3]) { return read3(ctx, ctx ? ctx->ans : nullptr, w); }


int clb_normal(clb_ctx* ctx, uint64_t n, uint64_t frac[3], char hex[49], uint64_t* n_out) {
  if (int rc = check_ctx(ctx)) return rc;
  uint64_t d, nn;
  if (int rc = normalize(n, &nn, &d, nullptr)) return rc;
  std::vector<HostJob> jobs;
  std::vector<uint8_t> shift(7);
  for (int j = 0; j < 7; ++j) {
    jobs.push_back(batch_job(d, 0, d, kDen0[j], kDen1[j], (uint32_t)kNumeSign[j], (uint32_t)j));
    shift[j] = kShift[j];
  }
  std::vector<GroupDev> groups{{0, 7}};
  uint64_t res[3];
  // Leave the device state as Allloop does: dev_ans = the answer, bigSum = formula 6's partial.
  Mirrors mirrors;
  mirrors.total = ctx->ans;
  mirrors.slot = ctx->partial;
  mirrors.index = 6;
  if (int rc = run(ctx, jobs, shift, groups, res, nullptr, nullptr, mirrors)) return rc;
  if (frac) memcpy(frac, res, sizeof res);
  if (hex) format_result(out_format(ctx, ctx->precision), res, hex);
  if (n_out) *n_out = nn;
  return CLB_OK;
}

int clb_batch_positions(clb_ctx* ctx, const uint64_t* n, size_t count, uint64_t* frac) {
  if (int rc = check_ctx(ctx)) return rc;
  if (count == 0) return CLB_OK;
  if (!n || !frac) return fail(CLB_EINVAL, "null array");
  // About 1.2 KB of the call's block per position (7 jobs, 7 tails, accumulators, result), which
  // must stay below 4 GiB.
  if (count > (1u << 21)) return fail(CLB_EINVAL, "too many positions (at most 2^21 per call)");
  std::vector<HostJob> jobs;
  std::vector<uint8_t> shift(7 * count);
  std::vector<GroupDev> groups(count);
  jobs.reserve(7 * count);
  for (size_t i = 0; i < count; ++i) {
    uint64_t d;
    if (int rc = normalize(n[i], nullptr, &d, nullptr)) return rc;
    for (int j = 0; j < 7; ++j) {
      const uint32_t slot = (uint32_t)(7 * i + j);
      jobs.push_back(batch_job(d, 0, d, kDen0[j], kDen1[j], (uint32_t)kNumeSign[j], slot));
      shift[slot] = kShift[j];
    }
    groups[i] = {(uint32_t)(7 * i), 7};
  }
  return run(ctx, jobs, shift, groups, frac, nullptr, nullptr);
}

int clb_run_jobs(clb_ctx* ctx, const clb_job* jobs_in, size_t count, uint64_t* partial) {
  if (int rc = check_ctx(ctx)) return rc;
  if (count == 0) return CLB_OK;
  if (!jobs_in || !partial) return fail(CLB_EINVAL, "null array");
  // About 200 bytes of the call's block per job, which must stay below 4 GiB.
  if (count > (1u << 24)) return fail(CLB_EINVAL, "too many jobs (at most 2^24 per call)");
  std::vector<HostJob> jobs;
  std::vector<uint8_t> shift(count);
  std::vector<GroupDev> groups(count);
  for (size_t i = 0; i < count; ++i) {
    const clb_job& jb = jobs_in[i];
    if (jb.formula < 0 || jb.formula > 6)
      return fail(CLB_EINVAL, "job %zu: formula %d", i, jb.formula);
    if (jb.d == 0 || jb.d > kMaxDigits) return fail(CLB_EINVAL, "job %zu: bad d", i);
    const int j = jb.formula;
    jobs.push_back(batch_job(jb.d, jb.k_begin, jb.k_end, kDen0[j], kDen1[j],
                             (uint32_t)kNumeSign[j], (uint32_t)i));
    shift[i] = kShift[j];
    groups[i] = {(uint32_t)i, 1};
  }
  return run(ctx, jobs, shift, groups, partial, nullptr, nullptr);
}

// clb_run_jobs without allocating: the call's block is built in the device's pinned staging
// buffer and copied into its file-scope device array (clb_noalloc.hip), then the same kernels run
// on it.  No heap, no device or pinned allocation, no graph (direct launches).
int clb_run_jobs_noalloc(clb_ctx* ctx, const clb_job* jobs_in, size_t count, uint64_t* partial) {
  if (int rc = check_ctx(ctx)) return rc;
  if (count == 0) return CLB_OK;
  if (!jobs_in || !partial) return fail(CLB_EINVAL, "null array");
  if (count > kNoAllocMaxJobs)
    return fail(CLB_EINVAL, "%zu jobs: the no-allocation entry point takes at most %u per call",
                count, kNoAllocMaxJobs);
  if (ctx->replay_gws)
    return fail(CLB_EINVAL, "the no-allocation entry point does not replay the reference");
  const auto t_call = std::chrono::steady_clock::now();
  uint8_t *dev = nullptr, *stage = nullptr;
  std::mutex* call_lock = nullptr;
  if (ctx->noalloc_err != hipSuccess || !noalloc_area(ctx->device, &dev, &stage, &call_lock))
    return fail(CLB_EDEVICE, "no-allocation area of device %d not set up: %s", ctx->device,
                hipGetErrorString(ctx->noalloc_err));
  if (int rc = set_device(ctx)) return rc;
  auto job_of = [&](size_t i) {
    const int j = jobs_in[i].formula;
    return batch_job(jobs_in[i].d, jobs_in[i].k_begin, jobs_in[i].k_end, kDen0[j], kDen1[j],
                     (uint32_t)kNumeSign[j], (uint32_t)i);
  };
  uint64_t max_m = 1;
  size_t nhead = 0, ntails = 0;
  for (size_t i = 0; i < count; ++i) {
    const clb_job& jb = jobs_in[i];
    if (jb.formula < 0 || jb.formula > 6)
      return fail(CLB_EINVAL, "job %zu: formula %d", i, jb.formula);
    if (jb.d == 0 || jb.d > kMaxDigits) return fail(CLB_EINVAL, "job %zu: bad d", i);
    const HostJob h = job_of(i);
    if (h.tail) {
      if (h.d + kTailTerms > ((1ull << 63) - h.den1) / h.den0)
        return fail(CLB_ERANGE, "tail modulus of d = %llu exceeds 2^63", (unsigned long long)h.d);
      ++ntails;
    }
    if (h.k0 >= h.k1) continue;
    const uint64_t m = h.den0 * (h.k1 - 1) + h.den1;
    if (m > max_m) max_m = m;
    ++nhead;
  }
  int kind = kSplitW3;
  if (nhead)
    if (int rc = choose_kind(ctx->kernel, max_m, &kind)) return rc;
  LaunchPlan plan;
  plan_launch(ctx, kind, count, [&](uint64_t tk) {
    uint64_t f = 0;
    for (size_t i = 0; i < count; ++i) {
      const HostJob h = job_of(i);
      if (h.k0 < h.k1) f += (h.k1 + tk - 1) / tk - h.k0 / tk;
    }
    return f;
  }, &plan);
  auto align64 = [](size_t x) { return (x + 63) & ~(size_t)63; };
  const size_t off_jobs = align64(sizeof(LaunchDev));
  const size_t off_tails = off_jobs + align64(nhead * sizeof(JobDev));
  const size_t off_groups = off_tails + align64(ntails * sizeof(TailDev));
  const size_t off_shifts = off_groups + align64(count * sizeof(GroupDev));
  const size_t off_acc = off_shifts + align64(count);
  const size_t off_out = off_acc + count * plan.replicas * kAccWords * sizeof(uint64_t);
  const size_t bytes = off_out + 3 * count * sizeof(uint64_t);
  if (bytes > kNoAllocBytes)
    return fail(CLB_EINVAL, "the call's block (%zu bytes) exceeds the no-allocation area (%zu)",
                bytes, kNoAllocBytes);
  std::lock_guard<std::mutex> lock(*call_lock);  // one call at a time uses this device's area
  LaunchDev h{};
  h.total_tiles = plan.flat;
  plan_to_header(plan, &h);
  h.dst = (uint64_t)(uintptr_t)(dev + off_out);
  h.njobs = (uint32_t)nhead;
  h.ntails = (uint32_t)ntails;
  h.ngroups = (uint32_t)count;
  h.nslots = (uint32_t)count;
  h.off_jobs = (uint32_t)off_jobs;
  h.off_tails = (uint32_t)off_tails;
  h.off_groups = (uint32_t)off_groups;
  h.off_shifts = (uint32_t)off_shifts;
  h.off_acc = (uint32_t)off_acc;
  h.off_out = (uint32_t)off_out;
  h.off_rjobs = h.off_rtails = (uint32_t)off_groups;  // no replay tables
  memcpy(stage, &h, sizeof h);
  const uint64_t tk = tile_k(kTplChoices[plan.tpl]);
  JobDev* jd = reinterpret_cast<JobDev*>(stage + off_jobs);
  TailDev* td = reinterpret_cast<TailDev*>(stage + off_tails);
  GroupDev* gd = reinterpret_cast<GroupDev*>(stage + off_groups);
  uint64_t flat = 0;
  for (size_t i = 0; i < count; ++i) {
    const HostJob hj = job_of(i);
    gd[i] = {(uint32_t)i, 1};
    stage[off_shifts + i] = kShift[jobs_in[i].formula];
    if (hj.tail) *td++ = {hj.d, hj.den0, hj.den1, hj.numesign, hj.slot};
    if (hj.k0 >= hj.k1) continue;
    JobDev j{};
    j.d = hj.d;
    j.k0 = hj.k0;
    j.k1 = hj.k1;
    j.tile_first = hj.k0 / tk;
    j.flat_begin = flat;
    j.den0 = hj.den0;
    j.den1 = hj.den1;
    j.numesign = hj.numesign;
    j.slot = hj.slot;
    flat += (hj.k1 + tk - 1) / tk - j.tile_first;
    *jd++ = j;
  }
  memset(stage + off_acc, 0, off_out - off_acc);  // the accumulator starts at zero every call
  LaunchDev* hdr = reinterpret_cast<LaunchDev*>(dev);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dev, stage, off_out, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ctx->ev[0], s));
  HIP_TRY(launch_head(kind, plan.tpl, hdr, (int)plan.grid, plan.wg_waves, s));
  HIP_TRY(hipEventRecord(ctx->ev[1], s));
  HIP_TRY(launch_combine(hdr, s));
  HIP_TRY(hipMemcpyAsync(stage + off_out, dev + off_out, 3 * count * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, s));
  hipError_t e2 = hipStreamSynchronize(s);
  float head = 0.f;
  if (e2 == hipSuccess) e2 = hipEventElapsedTime(&head, ctx->ev[0], ctx->ev[1]);
  if (e2 != hipSuccess) return fail(CLB_EDEVICE, "term launch failed: %s", hipGetErrorString(e2));
  memcpy(partial, stage + off_out, 3 * count * sizeof(uint64_t));
  ctx->head_ms = head;
  ctx->clk_valid = false;  // its header carries no clock address
  ctx->total_ms = (float)std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t_call)
                      .count();
  return CLB_OK;
}

int clb_last_timing(clb_ctx* ctx, double* head_kernel_ms, double* total_ms) {
  if (int rc = check_ctx(ctx)) return rc;
  if (head_kernel_ms) *head_kernel_ms = ctx->head_ms;
  if (total_ms) *total_ms = ctx->total_ms;
  return CLB_OK;
}

static_assert(CLB_CLOCK_XCDS == kClockXcds, "clock samples: one per XCD");

int clb_last_clock(clb_ctx* ctx, double ghz[CLB_CLOCK_XCDS], double* mean_ghz) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->clk_valid)
    return fail(CLB_EINVAL, "the last call ran no head-term kernel with clock samples");
  if (int rc = set_device(ctx)) return rc;
  uint64_t w[4 * kClockXcds];
  HIP_TRY(hipMemcpyAsync(w, ctx->clk, sizeof w, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  double sum = 0.0;
  int n = 0;
  for (int x = 0; x < kClockXcds; ++x) {
    const uint64_t cycles = w[4 * x + 2] - w[4 * x], ticks = w[4 * x + 3] - w[4 * x + 1];
    // the real-time counter runs at 100 MHz: ticks / 100 microseconds
    const double g = ticks ? (double)cycles / ((double)ticks * 10.0) : 0.0;
    if (ghz) ghz[x] = g;
    if (ticks) {
      sum += g;
      ++n;
    }
  }
  if (mean_ghz) *mean_ghz = n ? sum / n : 0.0;
  return CLB_OK;
}

int clb_run_worklist(clb_ctx* ctx, const char* worklist_path, const char* work_dir,
                     uint64_t batch_size, const char* device_label, int* jobs_run) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!worklist_path || !work_dir) return fail(CLB_EINVAL, "null path");
  if (jobs_run) *jobs_run = 0;
  Worklist wl;
  if (int rc = parse_worklist(worklist_path, &wl)) return rc;
  if (wl.digits < 5) return fail(CLB_EFORMAT, "worklist digits must be >= 5");
  if (wl.precision < 0) return fail(CLB_EFORMAT, "worklist precision must be 0, 1 or 2");
  const int precision = wl.precision;
  uint64_t d;
  if (int rc = normalize((uint64_t)wl.digits, nullptr, &d, nullptr)) return rc;
  const uint64_t batch = batch_size ? batch_size : kBatchSize;
  // *Batchloop (:298-340)
  std::vector<clb_job> todo;
  std::vector<std::string> paths;
  std::vector<int> formulas;
  for (const auto& jb : wl.jobs) {
    const int j = jb.first;
    const int64_t b = jb.second;
    if ((uint64_t)b > UINT64_MAX / batch - 1)
      return fail(CLB_EFORMAT, "batch %lld too large", (long long)b);
    const uint64_t k_start = batch * (uint64_t)b, k_end = batch * ((uint64_t)b + 1);
    if (k_start > d) continue;  // :306
    char name[64];
    snprintf(name, sizeof name, "/%d/%d_%lld.txt", j, j, (long long)b);
    const std::string path = std::string(work_dir) + name;
    if (file_exists(path)) continue;  // :307-308, the resume rule
    bool dup = false;
    for (const auto& p : paths) dup = dup || p == path;
    if (dup) continue;
    todo.push_back({d, k_start, k_end, j, 0});
    paths.push_back(path);
    formulas.push_back(j);
  }
  if (todo.empty()) return CLB_OK;
  const std::string label =
      device_label ? std::string(device_label)
                   : std::to_string(ctx->device) + "\t" + ctx->device_name;
  // The reference runs and saves one job at a time (:309-333), so a crash loses one job.  Here
  // consecutive jobs share a persistent launch up to the context's bounds (an estimated duration,
  // a job count), and each launch's files are written before the next launch starts; each file's
  // third line is its job's share of the launch time, by estimated cost.
  std::vector<double> cost(todo.size());
  for (size_t i = 0; i < todo.size(); ++i)
    cost[i] = job_squares(d, todo[i].k_begin, todo[i].k_end);
  std::vector<uint64_t> part;
  uint32_t launches = 0;
  for (size_t first = 0; first < todo.size();) {
    if (ctx->wl_max_launches && launches == ctx->wl_max_launches) break;
    size_t end = first + 1;
    double est = cost[first];
    while (end < todo.size() && (!ctx->wl_max_jobs || end - first < ctx->wl_max_jobs) &&
           (ctx->wl_max_seconds <= 0.0 || (est + cost[end]) / ctx->wl_rate <= ctx->wl_max_seconds)) {
      est += cost[end];
      ++end;
    }
    part.assign(3 * (end - first), 0);
    const int saved_precision = ctx->precision;
    ctx->precision = precision;
    int rc = clb_run_jobs(ctx, todo.data() + first, end - first, part.data());
    ctx->precision = saved_precision;
    if (rc) return rc;
    ++launches;
    if (ctx->head_ms > 0.f) {  // refine the rate estimate for the next launch's bound
      const double r = est / (ctx->head_ms * 1e-3);
      if (r > 1e9) ctx->wl_rate = r;
    }
    for (size_t i = first; i < end; ++i) {
      const int64_t b = (int64_t)(todo[i].k_begin / batch);
      const long long ms = (long long)std::llround(ctx->total_ms * cost[i] / est);
      if ((rc = write_work_file(work_dir, formulas[i], b, &part[3 * (i - first)],
                                out_format(ctx, precision), label, ms)))
        return rc;
      if (jobs_run) ++*jobs_run;
    }
    first = end;
  }
  return CLB_OK;
}

int clb_set_worklist_launch(clb_ctx* ctx, uint32_t max_jobs, double max_seconds,
                            uint32_t max_launches) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!(max_seconds >= 0.0) || max_seconds > 1e9)
    return fail(CLB_EINVAL, "max_seconds must be in [0, 1e9]");
  ctx->wl_max_jobs = max_jobs;
  ctx->wl_max_seconds = max_seconds;
  ctx->wl_max_launches = max_launches;
  return CLB_OK;
}

}  // extern "C"

// clb_peaks.hip -- issue-rate microbenchmarks of the VALU instructions the term kernels are built
// from, and of the two engines' whole modular squares, on gfx950.  Every lane runs kChains
// independent dependency chains of one instruction, unrolled kUnroll times per loop trip, on a
// grid of 8 waves per SIMD; the host times the launch with HIP events.  One wave per XCD also
// reads the shader-clock counter (s_memtime) and the 100 MHz real-time counter (s_memrealtime) at
// its start and end, which gives the clock the run actually had, so rates can be stated per CU per
// clock without assuming the nominal 2.4 GHz.
#include <hip/hip_runtime.h>

#include "../../include/clbellard.h"
#include "clb_device.hpp"
#include "clb_internal.hpp"
#include "clb_wave.hpp"

namespace clb {
namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 8;
constexpr int kPeakBlocksPerCu = 8;  // 256-thread workgroups: 32 waves per CU, 8 per SIMD

// One instruction of each kind on a chain; inline assembly so that the instruction is exactly the
// one named (the compiler would fold repeated roundings, for one).  ALU instructions only.
template <int OP>
struct Step;

template <>
struct Step<CLB_PEAK_FMA_F64> {
  __device__ static void run(double& x, double a, double b) {
    asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(x) : "v"(a), "v"(b));
  }
};
template <>
struct Step<CLB_PEAK_MUL_F64> {
  __device__ static void run(double& x, double a, double) {
    asm volatile("v_mul_f64 %0, %0, %1" : "+v"(x) : "v"(a));
  }
};
template <>
struct Step<CLB_PEAK_ADD_F64> {
  __device__ static void run(double& x, double a, double) {
    asm volatile("v_add_f64 %0, %0, %1" : "+v"(x) : "v"(a));
  }
};
template <>
struct Step<CLB_PEAK_RNDNE_F64> {
  __device__ static void run(double& x, double, double) {
    asm volatile("v_rndne_f64 %0, %0" : "+v"(x));
  }
};
template <>
struct Step<CLB_PEAK_LDEXP_F64> {
  __device__ static void run(double& x, double, double) {
    asm volatile("v_ldexp_f64 %0, %0, 0" : "+v"(x));
  }
};

template <int OP>
struct StepU;
template <>
struct StepU<CLB_PEAK_MAD_U64_U32> {
  __device__ static void run(uint64_t& x, uint32_t a, uint32_t b) {
    uint64_t carry;
    asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(x), "=s"(carry) : "v"(a), "v"(b));
  }
};
template <>
struct StepU<CLB_PEAK_MUL_HI_U32> {
  __device__ static void run(uint64_t& x, uint32_t a, uint32_t) {
    uint32_t v = (uint32_t)x;
    asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(v) : "v"(a));
    x = v;
  }
};
template <>
struct StepU<CLB_PEAK_MUL_LO_U32> {
  __device__ static void run(uint64_t& x, uint32_t a, uint32_t) {
    uint32_t v = (uint32_t)x;
    asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(v) : "v"(a));
    x = v;
  }
};
template <>
struct StepU<CLB_PEAK_ADD_U32> {
  __device__ static void run(uint64_t& x, uint32_t a, uint32_t) {
    uint32_t v = (uint32_t)x;
    asm volatile("v_add_u32 %0, %0, %1" : "+v"(v) : "v"(a));
    x = v;
  }
};
template <>
struct StepU<CLB_PEAK_FMA_F32> {
  __device__ static void run(uint64_t& x, uint32_t a, uint32_t b) {
    float v = __builtin_bit_cast(float, (uint32_t)x);
    asm volatile("v_fma_f32 %0, %0, %1, %2"
                 : "+v"(v)
                 : "v"(__builtin_bit_cast(float, a)), "v"(__builtin_bit_cast(float, b)));
    x = __builtin_bit_cast(uint32_t, v);
  }
};

// The first wave of each of the first kClockBlocks workgroups records both counters: workgroups
// are dispatched round-robin over the 8 XCDs, whose clocks may differ under load.
constexpr int kClockBlocks = 8;

__device__ __forceinline__ void clocks_begin(uint64_t* clk) {
  if (block_x() < kClockBlocks && thread_x() == 0) {
    clk[4 * block_x() + 0] = __builtin_amdgcn_s_memtime();
    clk[4 * block_x() + 1] = __builtin_amdgcn_s_memrealtime();
  }
}
__device__ __forceinline__ void clocks_end(uint64_t* clk) {
  if (block_x() < kClockBlocks && thread_x() == 0) {
    clk[4 * block_x() + 2] = __builtin_amdgcn_s_memtime();
    clk[4 * block_x() + 3] = __builtin_amdgcn_s_memrealtime();
  }
}

template <int OP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_peak_f64(uint64_t iters, uint64_t* clk, uint64_t* sink) {
  clocks_begin(clk);
  const double a = 1.0 + 0x1p-40 * (double)(thread_x() & 7), b = 0x1p-30;
  double x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = 1.0 + 0x1p-20 * c;
  for (uint64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int c = 0; c < kChains; ++c) Step<OP>::run(x[c], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (s == 0.125) sink[0] = 1;  // never true; keeps the chains live
  clocks_end(clk);
}

template <int OP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_peak_u32(uint64_t iters, uint64_t* clk, uint64_t* sink) {
  clocks_begin(clk);
  const uint32_t a = 0x9e3779b9u + thread_x(), b = 0x7f4a7c15u;
  uint64_t x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = 0x3f800000u + (uint64_t)c;
  for (uint64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int c = 0; c < kChains; ++c) StepU<OP>::run(x[c], a, b);
  }
  uint64_t s = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (s == 12345) sink[0] = s;
  clocks_end(clk);
}

// Whole modular squares of the two engines on chains of their own residues (m near 2^41).
constexpr int kSqChains = 4;

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_peak_fsq(uint64_t iters, uint64_t* clk, uint64_t* sink) {
  clocks_begin(clk);
  const double m = 2199023255531.0 + 2.0 * (double)thread_x();  // odd, below 2^41
  const double minv = recip(m);
  double x[kSqChains];
#pragma unroll
  for (int c = 0; c < kSqChains; ++c) x[c] = 3.0 + c;
  for (uint64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int c = 0; c < kSqChains; ++c) x[c] = fsq(x[c], m, minv);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kSqChains; ++c) s += x[c];
  if (s == 0.5) sink[0] = 1;
  clocks_end(clk);
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_peak_isq(uint64_t iters, uint64_t* clk, uint64_t* sink) {
  clocks_begin(clk);
  const uint64_t m = 2199023255531ull + 2ull * thread_x();
  const uint64_t mp = neg_inv64(m);
  uint64_t h[kSqChains];
#pragma unroll
  for (int c = 0; c < kSqChains; ++c) h[c] = 3 + c;
  for (uint64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int c = 0; c < kSqChains; ++c) h[c] = isq55(h[c], m, mp);
  }
  uint64_t s = 0;
#pragma unroll
  for (int c = 0; c < kSqChains; ++c) s += h[c];
  if (s == 12345) sink[0] = s;
  clocks_end(clk);
}

hipError_t launch_peak(int op, uint64_t iters, int grid, uint64_t* clk, uint64_t* sink,
                       hipStream_t s) {
  const dim3 g(grid), b(kBlock);
  switch (op) {
    case CLB_PEAK_FMA_F64: hipLaunchKernelGGL(k_peak_f64<CLB_PEAK_FMA_F64>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_MUL_F64: hipLaunchKernelGGL(k_peak_f64<CLB_PEAK_MUL_F64>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_ADD_F64: hipLaunchKernelGGL(k_peak_f64<CLB_PEAK_ADD_F64>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_RNDNE_F64: hipLaunchKernelGGL(k_peak_f64<CLB_PEAK_RNDNE_F64>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_LDEXP_F64: hipLaunchKernelGGL(k_peak_f64<CLB_PEAK_LDEXP_F64>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_MAD_U64_U32: hipLaunchKernelGGL(k_peak_u32<CLB_PEAK_MAD_U64_U32>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_MUL_HI_U32: hipLaunchKernelGGL(k_peak_u32<CLB_PEAK_MUL_HI_U32>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_MUL_LO_U32: hipLaunchKernelGGL(k_peak_u32<CLB_PEAK_MUL_LO_U32>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_ADD_U32: hipLaunchKernelGGL(k_peak_u32<CLB_PEAK_ADD_U32>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_FMA_F32: hipLaunchKernelGGL(k_peak_u32<CLB_PEAK_FMA_F32>, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_FSQ: hipLaunchKernelGGL(k_peak_fsq, g, b, 0, s, iters, clk, sink); break;
    case CLB_PEAK_ISQ: hipLaunchKernelGGL(k_peak_isq, g, b, 0, s, iters, clk, sink); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace clb

extern "C" int clb_issue_rate(int device, int op, double target_ms, double out[4]) {
  using namespace clb;
  if (!out || op < 0 || op >= CLB_PEAK_COUNT || !(target_ms > 0.0) || target_ms > 60000.0)
    return CLB_EINVAL;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess)
    return CLB_EDEVICE;
  const int grid = prop.multiProcessorCount * kPeakBlocksPerCu;
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  uint64_t* buf = nullptr;  // clocks [0, 4 kClockBlocks), sink after them
  uint64_t clk[4 * kClockBlocks] = {};
  int rc = CLB_EDEVICE;
  float ms = 0.f;
  uint64_t iters = 64;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) goto done;
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) goto done;
  if (hipMalloc((void**)&buf, (4 * kClockBlocks + 4) * sizeof(uint64_t)) != hipSuccess) goto done;
  // probe, then one run sized to the target time
  for (int pass = 0; pass < 2; ++pass) {
    if (hipEventRecord(ev[0], s) != hipSuccess || launch_peak(op, iters, grid, buf, buf + 4 * kClockBlocks, s) ||
        hipEventRecord(ev[1], s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess ||
        hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess)
      goto done;
    if (pass == 0) {
      const double scale = target_ms / (ms > 0.01f ? ms : 0.01f);
      iters = (uint64_t)((double)iters * scale);
      if (iters < 64) iters = 64;
    }
  }
  if (hipMemcpy(clk, buf, sizeof clk, hipMemcpyDeviceToHost) != hipSuccess) goto done;
  {
    const bool sq = op == CLB_PEAK_FSQ || op == CLB_PEAK_ISQ;
    const double per_lane = (double)iters * kUnroll * (sq ? kSqChains : kChains);
    const double lanes = (double)grid * kBlock;
    out[0] = per_lane * lanes / (ms * 1e-3);  // lane-operations (or modular squares) per second
    double ghz = 0.0;  // mean shader clock of the sampled waves
    for (int b = 0; b < kClockBlocks; ++b) {
      const uint64_t* c = clk + 4 * b;
      const double rt = (double)(c[3] - c[1]) / 100e6;  // the wave's seconds
      ghz += rt > 0.0 ? (double)(c[2] - c[0]) / rt / 1e9 / kClockBlocks : 0.0;
    }
    out[1] = ghz;
    out[2] = ms;
    out[3] = (double)prop.multiProcessorCount;
  }
  rc = CLB_OK;
done:
  if (buf) (void)hipFree(buf);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

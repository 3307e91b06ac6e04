// clb_wave.hpp -- work-item indices, wave-level 192-bit reductions and the accumulator's atomic
// limb adds, shared by the term kernels (clb_kernels.hip) and the reference-replay kernels
// (clb_replay.hip).
#pragma once

#include "clb_device.hpp"
#include "clb_internal.hpp"

namespace clb {

// Work-item and work-group indices from the hardware registers, block size and count from the
// hidden kernel arguments (code object v5: hidden_block_count_x at offset 0, hidden_group_size_x
// at offset 12, both in the kernarg segment in device memory).  threadIdx and friends go through
// device-library functions built for the default FP mode, which cannot be inlined into code built
// with the IEEE mode bit clear, and the calls would cost registers and spills.  The grid size is
// not taken from the dispatch packet: that lives in the queue's host-memory ring buffer, and every
// wave reading it added up to 13 us to a small launch.
__device__ __forceinline__ uint32_t thread_x() { return __builtin_amdgcn_workitem_id_x(); }
__device__ __forceinline__ uint32_t block_x() { return __builtin_amdgcn_workgroup_id_x(); }
__device__ __forceinline__ uint32_t block_y() { return __builtin_amdgcn_workgroup_id_y(); }
__device__ __forceinline__ uint32_t block_dim_x() { return __builtin_amdgcn_workgroup_size_x(); }
__device__ __forceinline__ uint32_t grid_dim_x() {
  return *(const uint32_t*)__builtin_amdgcn_implicitarg_ptr();  // C cast: address space 4
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int mask) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, mask, kWave);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), mask, kWave);
  return ((uint64_t)hi << 32) | lo;
}

// Sum of a 192-bit value over the 64 lanes of the wave (every lane receives the sum).
__device__ __forceinline__ void wave_sum192(U192& v) {
#pragma unroll
  for (int mask = 32; mask >= 1; mask >>= 1) {
    const uint64_t o0 = shfl_xor64(v.w0, mask);
    const uint64_t o1 = shfl_xor64(v.w1, mask);
    const uint64_t o2 = shfl_xor64(v.w2, mask);
    add192(v, o0, o1, o2);
  }
}

__device__ __forceinline__ void atomic_add192(unsigned long long* slot, const U192& v) {
  atomicAdd(slot + 0, (unsigned long long)(v.w0 & 0xffffffffull));
  atomicAdd(slot + 1, (unsigned long long)(v.w0 >> 32));
  atomicAdd(slot + 2, (unsigned long long)(v.w1 & 0xffffffffull));
  atomicAdd(slot + 3, (unsigned long long)(v.w1 >> 32));
  atomicAdd(slot + 4, (unsigned long long)(v.w2 & 0xffffffffull));
  atomicAdd(slot + 5, (unsigned long long)(v.w2 >> 32));
}

}  // namespace clb

// clb_noalloc.hip -- device memory of the no-allocation entry point (clb_run_jobs_noalloc): one
// file-scope array per GPU holding a call's block (header, job and tail tables, groups, shifts,
// accumulator, results), and a pinned host staging buffer of the same size, both set up once per
// device when the first context on it is created and kept until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>

#include "clb_internal.hpp"

namespace clb {

// The array itself.  In host code its name is only a stand-in for the symbol: every host access
// goes through the device address that hipGetSymbolAddress returns on each GPU (noalloc_setup).
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

struct Area {
  std::mutex call;          // held for the whole of one call on this device
  bool ready = false;
  uint8_t* dev = nullptr;   // this device's copy of `scratch`
  uint8_t* host = nullptr;  // pinned staging for the block
};

Area g_area[kMaxDevices];
std::mutex g_setup;

}  // namespace

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_setup);
  Area& a = g_area[device];
  if (a.ready) return hipSuccess;
  // `device` is current (clb_ctx_create calls this after hipSetDevice): the symbol's address on it
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  void* h = nullptr;
  if ((e = hipHostMalloc(&h, kNoAllocBytes, hipHostMallocDefault)) != hipSuccess) return e;
  a.dev = static_cast<uint8_t*>(p);
  a.host = static_cast<uint8_t*>(h);
  a.ready = true;
  return hipSuccess;
}

bool noalloc_area(int device, uint8_t** dev, uint8_t** host, std::mutex** call) {
  if (device < 0 || device >= kMaxDevices) return false;
  std::lock_guard<std::mutex> lock(g_setup);
  const Area& a = g_area[device];
  if (!a.ready) return false;
  *dev = a.dev;
  *host = a.host;
  *call = &g_area[device].call;
  return true;
}

}  // namespace clb

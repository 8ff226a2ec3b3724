// Device scratch of the allocation-free evolution (dq_sv_evolve_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) hand over the state
// buffer themselves; the only other device memory the evolution needs is the
// plan's constants (gate matrices, MFMA embeddings, tile-sweep blobs), which
// live in this file-scope array: kNoAllocScratchBytes on every GPU the library
// runs on, held until the process ends. Every call on a GPU shares that GPU's
// array and overwrites it; calls that overlap on one GPU race.
//
// In host code `scratch` names a host-side stand-in, not a device address:
// its device address is looked up once per device, with that device current,
// and only that address is ever passed to kernels or copies.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/dqsim_sv.h"
#include "engine.hpp"

namespace dq {
namespace {

constexpr uint64_t kNoAllocScratchBytes = DQ_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

void* noalloc_scratch(uint64_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw DeviceErr(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw DeviceErr("device index out of range for the allocation-free scratch");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw DeviceErr(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_addr[dev] = p;
  }
  *bytes = kNoAllocScratchBytes;
  return g_addr[dev];
}

}  // namespace dq

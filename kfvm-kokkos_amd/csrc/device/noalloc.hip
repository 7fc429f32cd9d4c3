// Scratch memory of the entry point that allocates nothing
// (kfvm_eval_rhs_noalloc): one file-scope device array per GPU, sized
// kNoAllocScratchBytes, that the HIP runtime places with the code object.
//
// In host code the name `scratch` is only a host-side stand-in for the
// symbol: its device address comes from hipGetSymbolAddress, taken once per
// device with that device current and kept here, and every kernel argument,
// fill or copy uses that address.  Every call on one GPU shares the array;
// calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "device_api.hpp"

namespace kfvm {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};
}  // namespace

void *noalloc_scratch() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_addr[dev] = p;
  }
  return g_addr[dev];
}

}  // namespace kfvm

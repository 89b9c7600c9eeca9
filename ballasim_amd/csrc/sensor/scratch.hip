// The library's static device scratch, for the entry point that allocates nothing
// (SensorModel::advance_device_noalloc): a fixed array in the code object, present on every GPU the
// library's kernels are loaded on until the process ends. Its device address differs per GPU and is
// taken once per GPU with hipGetSymbolAddress (the array's name in host code is only a stand-in for
// the symbol, never an address).
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "scratch.h"

namespace bsx {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void* scratch_address() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("bsx scratch: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("bsx scratch: device index beyond 63");
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      throw std::runtime_error(std::string("bsx scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    }
    addr[dev] = p;
  }
  return addr[dev];
}

}  // namespace bsx

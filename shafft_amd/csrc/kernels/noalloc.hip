// Scratch memory of the allocation-free FFT1D execute
// (shafftAmd1DExecuteNoAlloc): one file-scope device array, resolved to a
// device address once per device and then shared by every call on that
// device.  Calls that overlap on one device race on it; nothing in it has
// to start at zero (every pass writes the intermediate it leaves there
// before a later pass reads it, zero padding included).
#include <hip/hip_runtime.h>

#include <mutex>

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

// Device address and size of the array on the current device.  The first
// call on a device resolves the symbol with that device current; the name
// `scratch` itself is only a host-side stand-in and is never passed on.
extern "C" int sa_noalloc_scratch(void** ptr, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  *bytes = kNoAllocScratchBytes;
  return 0;
}

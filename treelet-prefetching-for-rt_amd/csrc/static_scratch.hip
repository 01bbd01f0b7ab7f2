// static_scratch.hip — the device memory rtx_trace_static works in.
//
// rtx_trace_static allocates nothing: its control words (ray counter, error
// word) and the traversal stacks that overflow LDS live in one file-scope
// device array.  Each GPU the library runs on holds its own copy of the
// array, at its own address, until the process ends.  In host code the
// array's name is not a device address, so its address is looked up once per
// device with hipGetSymbolAddress (with that device current) and only that
// address is handed to kernels and copies.
#include <mutex>

#include "kernels.h"

namespace rtxk {

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t static_scratch_init(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
  }
  *out = g_ptr[device];
  return hipSuccess;
}

}  // namespace rtxk

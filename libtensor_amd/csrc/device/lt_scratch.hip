// Fixed device scratch region of the no-allocation entry points.
//
// `scratch` is a file-scope device array: every GPU the library runs on
// holds its own copy, at its own address, from module load to process end.
// Host code never uses the array's name as a device address (in host code
// it names a host-side stand-in): noalloc_scratch() resolves the address
// once per device with hipGetSymbolAddress, with that device current, and
// every later use goes through the resolved pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "lt_device.h"

namespace {

static __device__ __align__(256) char scratch[lta::dev::kNoallocScratchBytes];

std::mutex g_mtx;
void *g_ptr[lta::dev::kMaxDevices];

} // unnamed namespace

namespace lta {
namespace dev {

hipError_t noalloc_scratch(int device, void **out) {
    *out = 0;
    if(device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mtx);
    if(!g_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if(e != hipSuccess) return e;
        if(cur != device) return hipErrorInvalidDevice;   // caller makes it current
        void *p = 0;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess) return e;
        if(!p) return hipErrorInvalidSymbol;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

} // namespace dev
} // namespace lta

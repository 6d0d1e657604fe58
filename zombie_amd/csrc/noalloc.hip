// noalloc.hip -- the fixed device workspace behind zg_wos_solve_noalloc.
//
// zg_wos_solve takes its per-call workspace (per-point FP64 sums and squares, walk counts and
// walk-length sums, 32 B per point, plus the solve's counters) from hipMallocAsync.  Callers
// that cannot allocate use zg_wos_solve_noalloc instead, whose workspace is the file-scope
// device array `scratch` below: each GPU the library runs on holds its own copy, from module
// load until the process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a host-side stand-in, not a device address); it asks noallocScratch() for
// the device address, looked up once per device with hipGetSymbolAddress while that device is
// current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace zg {

static __device__ __align__(256) unsigned long long scratch[kNoallocScratchBytes / 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
unsigned long long* g_scratchPtr[kMaxDevices] = {};
}  // namespace

hipError_t noallocScratch(unsigned long long** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratchPtr[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratchPtr[dev];
    return hipSuccess;
}

}  // namespace zg

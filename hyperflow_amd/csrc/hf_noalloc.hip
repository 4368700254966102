// hf_noalloc.hip -- the device scratch of hf_step_noalloc (include/hyperflow.h).
//
// hf_step_noalloc allocates no device memory: its tables (block records,
// problems, work items, the ghost gather table, Dirichlet constants) and its
// working arrays (wave-speed maxima, queue tickets, TimeStep call counters,
// speculation flags) live in one file-scope device array.  Each GPU holds its
// own copy, at its own address, until the process ends.  The array's name in
// host code is only a host-side stand-in; kernels receive the device address
// that hipGetSymbolAddress returns for the current device, looked up once per
// device here and cached.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hf_noalloc.h"

static __device__ __align__(256) unsigned char scratch[HF_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_mu;
void* g_addr[HF_NOALLOC_MAX_DEVICES] = {};
}  // namespace

int hf_noalloc_scratch(void** out, std::string* err)
{
    *out = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return -1;
    }
    if (dev < 0 || dev >= HF_NOALLOC_MAX_DEVICES) {
        *err = "device index " + std::to_string(dev) + " beyond the scratch table";
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return -1;
        }
        g_addr[dev] = p;
    }
    *out = g_addr[dev];
    return 0;
}

// Fixed device scratch for the entry points that allocate nothing
// (sqe_norm2_noalloc): a file-scope device array, one copy per GPU, whose
// device address is looked up once per device when an engine is created on
// it.  Host code passes that address to the kernels; the array's own name is
// only a host-side handle for the symbol lookup.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sqe_launch.hpp"

namespace sqe {

namespace {

static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return hipSuccess;
    int prev = 0;
    hipError_t r = hipGetDevice(&prev);
    if (r != hipSuccess) return r;
    if (prev != device && (r = hipSetDevice(device)) != hipSuccess) return r;
    void* p = nullptr;
    r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (r != hipSuccess) return r;
    g_ptr[device] = (double*)p;
    return hipSuccess;
}

double* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace sqe

// noalloc.hip -- entry gather for callers that cannot allocate device memory.
//
// dmx_get_entries stages the requested indices and the gathered entries in a
// device buffer it allocates per call.  Here that staging is one file-scope
// device array, `scratch`: every call on a device reuses that device's copy
// of it (indices first, then the real and the imaginary outputs).  Host code
// never passes the array's name to a kernel or a copy -- in host code that
// name is a host-side stand-in, not a device address -- only the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept.
#include "noalloc.hpp"

#include <mutex>

#include "sweep.hpp"

namespace dmx {

static __device__ __align__(256) unsigned char scratch[kNoAllocMaxEntries * 24];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void** out)
{
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

hipError_t gather_noalloc(const double* re, const double* im, int n, const uint64_t* idx,
                          uint64_t count, bool transposed, double* out_re, double* out_im,
                          hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    if (count > kNoAllocMaxEntries) return hipErrorInvalidValue;
    void* base = nullptr;
    hipError_t e = noalloc_scratch(&base);
    if (e != hipSuccess) return e;
    // every slot read below is written first in the same call: nothing needs
    // clearing between calls
    uint64_t* d_idx = static_cast<uint64_t*>(base);
    double* d_re = reinterpret_cast<double*>(d_idx + count);
    double* d_im = d_re + count;
    e = hipMemcpyAsync(d_idx, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = launch_gather(re, im, n, d_idx, count, transposed, d_re, d_im, stream);
    if (e == hipSuccess && out_re)
        e = hipMemcpyAsync(out_re, d_re, count * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && out_im)
        e = hipMemcpyAsync(out_im, d_im, count * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

}  // namespace dmx

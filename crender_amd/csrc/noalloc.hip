// noalloc.hip -- the device scratch of crt_intersect_noalloc (gfx950).
//
// crt_intersect stages rays and hits in buffers the context allocates on demand and traces
// with the context's spill stack; callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc) use
// crt_intersect_noalloc instead, whose scratch -- queue counters, leaf-order triangle and
// instance indices, barycentrics the caller does not want, and the traversal's per-lane
// spill stack for a fixed 64-block grid -- is this file-scope array.  Every GPU the library
// runs on holds its own copy (48 MiB) until the process ends; its address is looked up once
// per device (noalloc_scratch) and only that address is handed to kernels.
#include <mutex>

#include "crt_internal.h"

namespace crt {

namespace {

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mu;
void         *g_scratch[kMaxDevices] = {};

// count = n, work counters and error flag zero: a kernel rather than a memset, so the call
// also behaves under stream capture
__global__ void k_noalloc_init(uint32_t *cnt, uint32_t n)
{
    if (threadIdx.x < kNoAllocCounters) cnt[threadIdx.x] = threadIdx.x == 0 ? n : 0u;
}

}  // namespace

hipError_t noalloc_scratch(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if ((e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t e2 = hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (e2 != hipSuccess) return e2;
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return hipSuccess;
}

hipError_t launch_noalloc_init(uint32_t *cnt, uint32_t n, hipStream_t s)
{
    k_noalloc_init<<<1, kNoAllocCounters, 0, s>>>(cnt, n);
    return hipGetLastError();
}

}  // namespace crt

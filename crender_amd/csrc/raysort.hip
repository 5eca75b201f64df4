// raysort.hip -- reordering of an extension-ray queue for traversal coherence.
//
// Secondary rays leave their queue in path order: neighbouring rays start near each other
// but point anywhere, so a wave's lanes and an XCD's L2 see unrelated parts of the BVH.
// Sorting the queue by a key of (Morton code of the origin in the scene box, direction
// octant) before tracing makes each XCD's contiguous slice of the queue spatially
// coherent.  Only the order changes: every ray keeps its path, hits are written per queue
// entry, so results are identical with and without sorting.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "crt_internal.h"

namespace crt {

namespace {

__device__ __forceinline__ uint32_t spread3(uint32_t x)   // 9 bits -> every third bit
{
    x &= 0x1FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// key = morton27(origin) << 3 | octant(direction)
__global__ __launch_bounds__(256) void k_ray_keys(RaySoA q, const uint32_t *count, f3 bmin, f3 scale, uint32_t *keys,
                                                  uint32_t *vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const float fx = (q.ox[i] - bmin.x) * scale.x, fy = (q.oy[i] - bmin.y) * scale.y, fz = (q.oz[i] - bmin.z) * scale.z;
    const uint32_t x = (uint32_t)fminf(fmaxf(fx, 0.0f), 511.0f);
    const uint32_t y = (uint32_t)fminf(fmaxf(fy, 0.0f), 511.0f);
    const uint32_t z = (uint32_t)fminf(fmaxf(fz, 0.0f), 511.0f);
    const uint32_t oct = (q.dx[i] < 0.0f ? 1u : 0u) | (q.dy[i] < 0.0f ? 2u : 0u) | (q.dz[i] < 0.0f ? 4u : 0u);
    keys[i] = ((spread3(x) | (spread3(y) << 1) | (spread3(z) << 2)) << 3) | oct;
    vals[i] = i;
}

// dst[i] = src[perm[i]] for every ray component and the path id; dst count = src count
__global__ __launch_bounds__(256) void k_ray_gather(RaySoA src, const uint32_t *src_path, const uint32_t *count,
                                                    const uint32_t *perm, RaySoA dst, uint32_t *dst_path,
                                                    uint32_t *dst_count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *count;
    if (i == 0) *dst_count = n;
    if (i >= n) return;
    const uint32_t j = perm[i];
    dst.ox[i] = src.ox[j];
    dst.oy[i] = src.oy[j];
    dst.oz[i] = src.oz[j];
    dst.dx[i] = src.dx[j];
    dst.dy[i] = src.dy[j];
    dst.dz[i] = src.dz[j];
    dst.tn[i] = src.tn[j];
    dst.tf[i] = src.tf[j];
    dst_path[i] = src_path[j];
}

}  // namespace

size_t ray_sort_temp_bytes(uint32_t capacity)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)capacity, 0, 30);
    return bytes;
}

hipError_t launch_ray_sort(const RaySoA &src, const uint32_t *src_path, const uint32_t *src_count, uint32_t n,
                           const RaySoA &dst, uint32_t *dst_path, uint32_t *dst_count, f3 bmin, f3 bmax,
                           const RaySortScratch &w, hipStream_t s)
{
    const f3 ext   = bmax - bmin;
    const f3 scale = mk3(512.0f / fmaxf(ext.x, 1e-20f), 512.0f / fmaxf(ext.y, 1e-20f), 512.0f / fmaxf(ext.z, 1e-20f));
    if (n) {
        k_ray_keys<<<(n + 255) / 256, 256, 0, s>>>(src, src_count, bmin, scale, w.keys, w.vals);
        size_t bytes = w.temp_bytes;
        const hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.keys, w.keys_out, w.vals, w.vals_out,
                                                                (int)n, 0, 30, s);
        if (e != hipSuccess) return e;
    }
    k_ray_gather<<<n ? (n + 255) / 256 : 1, 256, 0, s>>>(src, src_path, src_count, w.vals_out, dst, dst_path,
                                                         dst_count);
    return hipGetLastError();
}

}  // namespace crt

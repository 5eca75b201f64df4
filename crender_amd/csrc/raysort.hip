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

// ---- material-sorted shading ----------------------------------------------------------
// Class of queue entry k: 0 = miss (sky), 1 + material type (metal 0, smooth 1, glass 2)
// otherwise, i.e. the branch of the shading code the entry takes.
__device__ __forceinline__ uint32_t shade_class(const SceneDev &sc, const HitSoA &h, uint32_t k)
{
    const uint32_t tri = h.tri[k];
    if (tri == kInvalidRef) return 0u;
    const uint32_t prim = __float_as_uint(sc.tri[3 * (size_t)tri].w);
    const int32_t  t    = sc.mats[sc.mat_idx[prim]].type;
    return 1u + (uint32_t)min(max(t, 0), 2);
}

// bcount[c * gridDim.x + block] = entries of class c among the block's 256 queue entries
__global__ __launch_bounds__(256) void k_class_count(SceneDev sc, HitSoA h, const uint32_t *count, uint32_t *bcount)
{
    __shared__ uint32_t wc[4][kShadeClasses];
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    const bool     in = k < *count;
    const uint32_t c  = in ? shade_class(sc, h, k) : kShadeClasses;
    const int      w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t j = 0; j < kShadeClasses; ++j) {
        const uint64_t m = __ballot(c == j);
        if (lane == 0) wc[w][j] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kShadeClasses) {
        const uint32_t j = threadIdx.x;
        bcount[j * gridDim.x + blockIdx.x] = wc[0][j] + wc[1][j] + wc[2][j] + wc[3][j];
    }
}

// In-place exclusive prefix sum of bcount[0 .. n) by one workgroup of 1024 lanes.
__global__ __launch_bounds__(1024) void k_class_scan(uint32_t *bcount, uint32_t n)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per, b1 = min(n, b0 + per);
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b1; ++i) sum += bcount[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {   // Hillis-Steele inclusive scan
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t v = bcount[i];
        bcount[i] = run;
        run += v;
    }
}

// perm[offset of (class, block) + rank within the block] = queue index: stable by index.
__global__ __launch_bounds__(256) void k_class_scatter(SceneDev sc, HitSoA h, const uint32_t *count,
                                                       const uint32_t *boff, uint32_t *perm)
{
    __shared__ uint32_t wc[4][kShadeClasses];
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    const bool     in = k < *count;
    const uint32_t c  = in ? shade_class(sc, h, k) : kShadeClasses;
    const int      w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t       rank = 0;
#pragma unroll
    for (uint32_t j = 0; j < kShadeClasses; ++j) {
        const uint64_t m = __ballot(c == j);
        if (lane == 0) wc[w][j] = (uint32_t)__popcll(m);
        if (c == j) rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (in) {
        uint32_t off = boff[c * gridDim.x + blockIdx.x] + rank;
        for (int v = 0; v < w; ++v) off += wc[v][c];
        perm[off] = k;
    }
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

hipError_t launch_shade_sort(const SceneDev &sc, const PathBuffers &b, uint32_t cur, hipStream_t s)
{
    const uint32_t nb = (b.P + 255) / 256;
    k_class_count<<<nb, 256, 0, s>>>(sc, b.hits, b.cnt + cur, b.shade_bcount);
    k_class_scan<<<1, 1024, 0, s>>>(b.shade_bcount, nb * kShadeClasses);
    k_class_scatter<<<nb, 256, 0, s>>>(sc, b.hits, b.cnt + cur, b.shade_bcount, b.shade_perm);
    return hipGetLastError();
}

}  // namespace crt

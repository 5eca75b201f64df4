// bvh_build.hip -- device BVH build for the crt path tracer (gfx950).
//
// Replaces the BVH that Embree builds inside rtcCommitScene (CRender
// src/objects/model.cpp:92-94).  Three stages, all on the device:
//   1. binned SAH (16 bins per axis over centroid bounds) top-down.  Segments larger than
//      one chunk (2048 primitives) are binned by many workgroups (LDS bins merged with
//      global atomics) and partitioned stably chunk by chunk; smaller segments are built
//      to completion by one workgroup each, with the primitive ids held in LDS.
//      Node slots are derived from segment ranges, so the tree is deterministic.
//   2. collapse of the binary tree into a 4- or 8-wide tree: an SAH-optimal distribution of
//      the binary subtrees over each wide node's child slots, small subtrees becoming
//      multi-primitive leaves (dynamic programme bottom-up, emission top-down level by level).
//   3. encoding into 128-byte nodes (FP32 boxes for 4-wide, 8-bit quantised boxes for
//      8-wide) and reordering of the triangles into leaf order.
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>
#include <functional>
#include <mutex>
#include <vector>
#include "crt_internal.h"

namespace crt {

namespace {

constexpr int kBinWords = 3 * kNumBins * 13;
constexpr float kInf = __builtin_huge_valf();

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float area3(float lx, float ly, float lz, float hx, float hy, float hz)
{
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Bin index of centroid coordinate c for a segment with centroid range [lo, hi]: the scale
// k = 16 (1 - 1e-6) / (hi - lo) is formed once per segment and axis (0 for an empty range)
// and bin_at applies it per primitive.  The binning and the partition both use these, so
// they agree bit for bit.
__device__ __forceinline__ float bin_scale(float lo, float hi)
{
    const float ext = hi - lo;
    return ext > 0.0f ? (float)kNumBins * (1.0f - 1e-6f) / ext : 0.0f;
}

__device__ __forceinline__ int bin_at(float c, float lo, float k)
{
    if (!(k > 0.0f)) return 0;
    int b = (int)((c - lo) * k);
    return b < 0 ? 0 : (b >= kNumBins ? kNumBins - 1 : b);
}

__device__ __forceinline__ int bin_of(float c, float lo, float hi) { return bin_at(c, lo, bin_scale(lo, hi)); }

__device__ __forceinline__ float4 scale4(const float *lo, const float *hi)
{
    return make_float4(bin_scale(lo[0], hi[0]), bin_scale(lo[1], hi[1]), bin_scale(lo[2], hi[2]), 0.0f);
}

// The centroid range is passed by value (its low corner and the per-axis scales): a pointer
// into a task held in registers would put the task in scratch memory.
__device__ __forceinline__ void bins_add(int *bins, float4 c, float4 lo, float4 hi, float4 cblo, float4 kscale)
{
    const float cc[3] = {c.x, c.y, c.z};
    const float bl[3] = {cblo.x, cblo.y, cblo.z}, ks[3] = {kscale.x, kscale.y, kscale.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = bin_at(cc[a], bl[a], ks[a]);
        int *p = bins + (a * kNumBins + b) * 13;
        atomicAdd(p + 0, 1);
        atomicMin(p + 1, f2ord(lo.x));
        atomicMin(p + 2, f2ord(lo.y));
        atomicMin(p + 3, f2ord(lo.z));
        atomicMax(p + 4, f2ord(hi.x));
        atomicMax(p + 5, f2ord(hi.y));
        atomicMax(p + 6, f2ord(hi.z));
        atomicMin(p + 7, f2ord(c.x));
        atomicMin(p + 8, f2ord(c.y));
        atomicMin(p + 9, f2ord(c.z));
        atomicMax(p + 10, f2ord(c.x));
        atomicMax(p + 11, f2ord(c.y));
        atomicMax(p + 12, f2ord(c.z));
    }
}

__device__ __forceinline__ float4 lo4(const float *v) { return make_float4(v[0], v[1], v[2], 0.0f); }
__device__ __forceinline__ float pick(float x, float y, float z, int a) { return a == 0 ? x : (a == 1 ? y : z); }

// One thread per (axis, bin) record of 13 words: count, box low / high, centroid low / high.
__device__ __forceinline__ void bins_init(int *bins, int tid, int nthreads)
{
    const int lo = f2ord(kInf), hi = f2ord(-kInf);
    for (int r = tid; r < 3 * kNumBins; r += nthreads) {
        int *p = bins + r * 13;
        p[0] = 0;
        p[1] = p[2] = p[3] = lo;
        p[4] = p[5] = p[6] = hi;
        p[7] = p[8] = p[9] = lo;
        p[10] = p[11] = p[12] = hi;
    }
}

struct BinVal {
    uint32_t n;   // exact count (a float sum loses integers above 2^24)
    float lo[3], hi[3], clo[3], chi[3];
};

__device__ __forceinline__ BinVal bv_identity()
{
    BinVal v;
    v.n = 0u;
    for (int k = 0; k < 3; ++k) {
        v.lo[k] = kInf;
        v.hi[k] = -kInf;
        v.clo[k] = kInf;
        v.chi[k] = -kInf;
    }
    return v;
}

__device__ __forceinline__ BinVal bv_join(const BinVal &a, const BinVal &b)
{
    BinVal r;
    r.n = a.n + b.n;
    for (int k = 0; k < 3; ++k) {
        r.lo[k]  = fminf(a.lo[k], b.lo[k]);
        r.hi[k]  = fmaxf(a.hi[k], b.hi[k]);
        r.clo[k] = fminf(a.clo[k], b.clo[k]);
        r.chi[k] = fmaxf(a.chi[k], b.chi[k]);
    }
    return r;
}

// Moves a value d lanes within its row of 16 lanes with a DPP row shift (a VALU operand
// modifier, no LDS traffic): kDir 0 takes lane - d (row_shr), 1 takes lane + d (row_shl).
// Lanes whose source falls outside the row keep their own value (callers ignore them).
template <int kDir, int d>
__device__ __forceinline__ uint32_t row_shift(uint32_t v)
{
    constexpr int ctrl = (kDir == 0 ? 0x110 : 0x100) + d;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, ctrl, 0xF, 0xF, false);
}
template <int kDir, int d>
__device__ __forceinline__ float row_shift(float v)
{
    return __uint_as_float(row_shift<kDir, d>(__float_as_uint(v)));
}

template <int kDir, int d>  // 0: from lane - d, 1: from lane + d, within rows of 16 lanes
__device__ __forceinline__ BinVal bv_shfl(const BinVal &v)
{
    BinVal r;
    r.n = row_shift<kDir, d>(v.n);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k]  = row_shift<kDir, d>(v.lo[k]);
        r.hi[k]  = row_shift<kDir, d>(v.hi[k]);
        r.clo[k] = row_shift<kDir, d>(v.clo[k]);
        r.chi[k] = row_shift<kDir, d>(v.chi[k]);
    }
    return r;
}

// Result of evaluating one segment's bins.
struct SahOut {
    float    lo[3], hi[3];        // node box
    uint32_t n;
    int32_t  axis, bin;           // best binned split (axis -1: none)
    uint32_t nleft;
    float    cost;                // AL*NL + AR*NR
    float    lcblo[3], lcbhi[3], rcblo[3], rcbhi[3];
};

// One wave evaluates all 3 x 15 candidate planes with segmented prefix/suffix scans.
// Must be called by all 64 lanes of one wave.
__device__ void sah_wave(const int *bins, SahOut *out)
{
    const int lane = threadIdx.x & 63;
    const int axis = lane >> 4, b = lane & 15;
    BinVal    v = bv_identity();
    if (lane < 48) {
        const int *p = bins + (axis * kNumBins + b) * 13;
        v.n = (uint32_t)p[0];
        if (p[0] > 0) {
            for (int k = 0; k < 3; ++k) {
                v.lo[k]  = ord2f(p[1 + k]);
                v.hi[k]  = ord2f(p[4 + k]);
                v.clo[k] = ord2f(p[7 + k]);
                v.chi[k] = ord2f(p[10 + k]);
            }
        }
    }
    // segmented (16-lane) inclusive prefix and suffix joins
    BinVal pre = v, suf = v;
#define SCAN_STEP(d)                                       \
    {                                                      \
        const BinVal o = bv_shfl<0, d>(pre);               \
        if (b >= d) pre = bv_join(pre, o);                 \
        const BinVal q = bv_shfl<1, d>(suf);               \
        if (b + d < 16) suf = bv_join(suf, q);             \
    }
    SCAN_STEP(1)
    SCAN_STEP(2)
    SCAN_STEP(4)
    SCAN_STEP(8)
#undef SCAN_STEP
    const BinVal rs = bv_shfl<1, 1>(suf);
    float cost = kInf;
    if (lane < 48 && b < 15 && pre.n > 0u && rs.n > 0u) {
        cost = area3(pre.lo[0], pre.lo[1], pre.lo[2], pre.hi[0], pre.hi[1], pre.hi[2]) * (float)pre.n +
               area3(rs.lo[0], rs.lo[1], rs.lo[2], rs.hi[0], rs.hi[1], rs.hi[2]) * (float)rs.n;
    }
    // argmin (cost, lane)
    float bc = cost;
    int   bl = lane;
    for (int o = 32; o > 0; o >>= 1) {
        const float oc = __shfl_xor(bc, o);
        const int   ol = __shfl_xor(bl, o);
        if (oc < bc || (oc == bc && ol < bl)) {
            bc = oc;
            bl = ol;
        }
    }
    // node box and count: full prefix of axis 0 (lane 15)
    const uint32_t tn = __shfl(pre.n, 15);
    float       nb[6];
    for (int k = 0; k < 3; ++k) {
        nb[k]     = __shfl(pre.lo[k], 15);
        nb[3 + k] = __shfl(pre.hi[k], 15);
    }
    if (lane == 0) {
        for (int k = 0; k < 3; ++k) {
            out->lo[k] = nb[k];
            out->hi[k] = nb[3 + k];
        }
        out->n = tn;
        out->cost = bc;
        if (bc == kInf) {
            out->axis = -1;
            out->bin  = -1;
        } else {
            out->axis = bl >> 4;
            out->bin  = bl & 15;
        }
    }
    if (bc != kInf && lane == bl) {
        out->nleft = pre.n;
        for (int k = 0; k < 3; ++k) {
            out->lcblo[k] = pre.clo[k];
            out->lcbhi[k] = pre.chi[k];
            out->rcblo[k] = rs.clo[k];
            out->rcbhi[k] = rs.chi[k];
        }
    }
}

// Leaf / split decision (C_trav = 0.5, C_isect = 1).  Returns 1 for a leaf; otherwise
// sets nleft and whether the split is by position (axis < 0).
__device__ __forceinline__ int decide(const SahOut &s, uint32_t max_leaf, uint32_t &nleft, int &axis)
{
    const uint32_t n = s.n;
    if (n <= 1) return 1;
    const float pa = area3(s.lo[0], s.lo[1], s.lo[2], s.hi[0], s.hi[1], s.hi[2]);
    if (s.axis >= 0 && pa > 0.0f) {
        const float split_cost = 0.5f + s.cost / pa;
        if (n <= max_leaf && (float)n <= split_cost) return 1;
        nleft = s.nleft;
        axis  = s.axis;
        return 0;
    }
    if (n <= max_leaf) return 1;
    nleft = n / 2;
    axis  = -1;
    return 0;
}

// Node writers.  A slot or child range outside the tree (nslots = 2 * ntri slots) flags the
// build as failed instead of writing out of bounds.
__device__ __forceinline__ void write_inner(Bvh2Node *nodes, uint32_t nslots, uint32_t *err, uint32_t slot,
                                            const SahOut &s, uint32_t nleft)
{
    if (slot >= nslots || nleft == 0u || nleft >= nslots || slot + 2u * nleft >= nslots) {
        atomicOr(err, 1u);
        return;
    }
    Bvh2Node nd;
    for (int k = 0; k < 3; ++k) {
        nd.lo[k] = s.lo[k];
        nd.hi[k] = s.hi[k];
    }
    nd.a        = slot + 1;
    nd.b        = slot + 2 * nleft;
    nodes[slot] = nd;
}

__device__ __forceinline__ void write_leaf(Bvh2Node *nodes, uint32_t nslots, uint32_t *err, uint32_t slot,
                                           const SahOut &s, uint32_t first, uint32_t n)
{
    if (slot >= nslots || n == 0u || n > 16u || first > nslots / 2u || n > nslots / 2u - first) {
        atomicOr(err, 1u);
        return;
    }
    Bvh2Node nd;
    for (int k = 0; k < 3; ++k) {
        nd.lo[k] = s.lo[k];
        nd.hi[k] = s.hi[k];
    }
    nd.a        = first;
    nd.b        = kLeafBit | n;
    nodes[slot] = nd;
}

// ---- stage 0: primitive boxes, centroids, root bounds ---------------------------------
__global__ __launch_bounds__(256) void k_prim_bounds(const float *__restrict__ tri, uint32_t n, float4 *plo,
                                                     float4 *phi, float4 *cen, uint32_t *idx, int *root)
{
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    float clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *p = tri + 9 * (size_t)i;
        float        l[3], h[3];
        for (int k = 0; k < 3; ++k) {
            l[k] = fminf(fminf(p[k], p[3 + k]), p[6 + k]);
            h[k] = fmaxf(fmaxf(p[k], p[3 + k]), p[6 + k]);
        }
        const float c[3] = {0.5f * (l[0] + h[0]), 0.5f * (l[1] + h[1]), 0.5f * (l[2] + h[2])};
        plo[i] = make_float4(l[0], l[1], l[2], 0.0f);
        phi[i] = make_float4(h[0], h[1], h[2], 0.0f);
        cen[i] = make_float4(c[0], c[1], c[2], 0.0f);
        idx[i] = i;
        for (int k = 0; k < 3; ++k) {
            lo[k]  = fminf(lo[k], l[k]);
            hi[k]  = fmaxf(hi[k], h[k]);
            clo[k] = fminf(clo[k], c[k]);
            chi[k] = fmaxf(chi[k], c[k]);
        }
    }
    __shared__ float red[4][12];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        lo[k]  = wave_min(lo[k]);
        hi[k]  = wave_max(hi[k]);
        clo[k] = wave_min(clo[k]);
        chi[k] = wave_max(chi[k]);
    }
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            red[w][k]     = lo[k];
            red[w][3 + k] = hi[k];
            red[w][6 + k] = clo[k];
            red[w][9 + k] = chi[k];
        }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        float     v = red[0][k];
        for (int j = 1; j < (int)(blockDim.x >> 6); ++j) v = ((k / 3) % 2 == 0) ? fminf(v, red[j][k]) : fmaxf(v, red[j][k]);
        if ((k / 3) % 2 == 0) atomicMin(root + k, f2ord(v));
        else atomicMax(root + k, f2ord(v));
    }
}

__global__ void k_root_task(const int *root, uint32_t n, BuildTask *large, BuildTask *small, uint32_t *nlarge,
                            uint32_t *counts)
{
    BuildTask t;
    t.begin = 0;
    t.end   = n;
    t.slot  = 0;
    t.pad   = 0;
    for (int k = 0; k < 3; ++k) {
        t.cblo[k] = ord2f(root[6 + k]);
        t.cbhi[k] = ord2f(root[9 + k]);
    }
    if (n > kChunk) {
        large[0]  = t;
        *nlarge   = 1;
        counts[1] = 0;  // small
    } else {
        small[0]  = t;
        *nlarge   = 0;
        counts[1] = 1;
    }
}

// ---- stage 1a: segments larger than one chunk --------------------------------------------
// Every kernel of a level reads its task and chunk counts from device memory and loops over
// them with a fixed grid, so the host launches several levels before it reads a count.
// Chunks of kChunk primitives of every large task, in task order: one block of 1024 threads,
// a scan of the per-task chunk counts, then one chunk per thread.  The task count is read
// from device memory (written by the previous level's split), so the host learns the next
// level's task and chunk counts with one read.
constexpr uint32_t kChunkBlock = 1024;

__global__ __launch_bounds__(kChunkBlock) void k_make_chunks(const BuildTask *tasks, const uint32_t *ntasks_p,
                                                             uint32_t *chunk_task, uint32_t *chunk_begin,
                                                             uint32_t *task_chunk0, uint32_t *nchunks)
{
    __shared__ uint32_t part[kChunkBlock];
    const uint32_t tid    = threadIdx.x;
    const uint32_t ntasks = *ntasks_p;
    const uint32_t per = (ntasks + kChunkBlock - 1) / kChunkBlock;
    const uint32_t t0 = min(tid * per, ntasks), t1 = min(t0 + per, ntasks);
    auto nch = [&](uint32_t t) { return (tasks[t].end - tasks[t].begin + kChunk - 1) / kChunk; };
    uint32_t local = 0;
    for (uint32_t t = t0; t < t1; ++t) local += nch(t);
    part[tid] = local;
    __syncthreads();
    for (uint32_t d = 1; d < kChunkBlock; d <<= 1) {   // inclusive scan
        const uint32_t v = tid >= d ? part[tid - d] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - local;
    for (uint32_t t = t0; t < t1; ++t) {
        task_chunk0[t] = run;
        run += nch(t);
    }
    const uint32_t total = part[kChunkBlock - 1];
    if (tid == 0) {
        task_chunk0[ntasks] = total;
        *nchunks            = total;
    }
    __syncthreads();   // task_chunk0 written by the block is visible to it
    for (uint32_t c = tid; c < total; c += kChunkBlock) {
        uint32_t lo = 0, hi = ntasks;   // the task t with task_chunk0[t] <= c < task_chunk0[t + 1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (task_chunk0[mid] <= c) lo = mid;
            else hi = mid;
        }
        chunk_task[c]  = lo;
        chunk_begin[c] = tasks[lo].begin + (c - task_chunk0[lo]) * kChunk;
    }
}

// Also zeroes the counter the level's split appends the next level's large tasks to.
__global__ void k_clear_bins(int *bins, const uint32_t *ntasks_p, uint32_t *next_count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0u;
    const uint32_t n = *ntasks_p;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) bins_init(bins + (size_t)t * kBinWords, threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(256) void k_bin_large(const BuildTask *tasks, const uint32_t *chunk_task,
                                                   const uint32_t *chunk_begin, const uint32_t *idx,
                                                   const float4 *plo, const float4 *phi, const float4 *cen,
                                                   int *gbins, const uint32_t *nchunks_p)
{
    __shared__ int bins[kBinWords];
    const uint32_t nch = *nchunks_p;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const BuildTask t = tasks[chunk_task[c]];
        const uint32_t  b = chunk_begin[c];
        const uint32_t  e = min(b + kChunk, t.end);
        bins_init(bins, threadIdx.x, blockDim.x);
        __syncthreads();
        const float4 cblo = lo4(t.cblo), ks = scale4(t.cblo, t.cbhi);
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
            const uint32_t id = idx[i];
            bins_add(bins, cen[id], plo[id], phi[id], cblo, ks);
        }
        __syncthreads();
        int *g = gbins + (size_t)chunk_task[c] * kBinWords;
        for (int i = threadIdx.x; i < kBinWords; i += blockDim.x) {
            const int w = i % 13;
            const int v = bins[i];
            if (w == 0) {
                if (v) atomicAdd(g + i, v);
            } else if ((w >= 1 && w <= 3) || (w >= 7 && w <= 9)) {
                atomicMin(g + i, v);
            } else {
                atomicMax(g + i, v);
            }
        }
        __syncthreads();   // bins are cleared for the next chunk
    }
}

// Split decision, node and children of one large segment (next_count: the next level's
// large-task counter, counts[1]: small tasks, counts[5]: error flag).
__device__ __forceinline__ void split_one(const SahOut &s, const BuildTask &tk, uint32_t t, SplitInfo *split,
                                          Bvh2Node *nodes, uint32_t nslots, BuildTask *next_large, BuildTask *small,
                                          uint32_t *next_count, uint32_t *counts, uint32_t max_leaf)
{
    uint32_t        nleft = 0;
    int             axis  = -1;
    SplitInfo       si;
    if (decide(s, max_leaf, nleft, axis)) {
        // cannot happen for n > kChunk >= max_leaf; keep a valid tree regardless
        nleft = s.n / 2;
        axis  = -1;
    }
    si.axis  = axis;
    si.bin   = axis >= 0 ? s.bin : -1;
    si.nleft = nleft;
    si.leaf  = 0;
    split[t] = si;
    write_inner(nodes, nslots, counts + 5, tk.slot, s, nleft);
    BuildTask l, r;
    l.begin = tk.begin;
    l.end   = tk.begin + nleft;
    l.slot  = tk.slot + 1;
    l.pad   = 0;
    r.begin = tk.begin + nleft;
    r.end   = tk.end;
    r.slot  = tk.slot + 2 * nleft;
    r.pad   = 0;
    for (int k = 0; k < 3; ++k) {
        l.cblo[k] = axis >= 0 ? s.lcblo[k] : tk.cblo[k];
        l.cbhi[k] = axis >= 0 ? s.lcbhi[k] : tk.cbhi[k];
        r.cblo[k] = axis >= 0 ? s.rcblo[k] : tk.cblo[k];
        r.cbhi[k] = axis >= 0 ? s.rcbhi[k] : tk.cbhi[k];
    }
    const BuildTask ch[2] = {l, r};
    for (int j = 0; j < 2; ++j) {
        if (ch[j].end - ch[j].begin > kChunk) next_large[atomicAdd(next_count, 1u)] = ch[j];
        else small[atomicAdd(counts + 1, 1u)] = ch[j];
    }
}

// one wave per large segment: choose the split, write the node, emit the children
__global__ __launch_bounds__(64) void k_split_large(const BuildTask *tasks, const uint32_t *ntasks_p,
                                                    const int *gbins, SplitInfo *split, Bvh2Node *nodes,
                                                    uint32_t nslots, BuildTask *next_large, BuildTask *small,
                                                    uint32_t *next_count, uint32_t *counts, uint32_t max_leaf)
{
    __shared__ SahOut s;
    const uint32_t n = *ntasks_p;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        sah_wave(gbins + (size_t)t * kBinWords, &s);
        __syncthreads();
        if (threadIdx.x == 0) split_one(s, tasks[t], t, split, nodes, nslots, next_large, small, next_count, counts,
                                        max_leaf);
        __syncthreads();   // s is rewritten for the next task
    }
}

__device__ __forceinline__ bool goes_left(const SplitInfo &si, const BuildTask &t, uint32_t pos, uint32_t id,
                                          const float4 *cen)
{
    if (si.axis < 0) return pos < t.begin + si.nleft;
    const float4 c  = cen[id];
    const float  cc = si.axis == 0 ? c.x : (si.axis == 1 ? c.y : c.z);
    return bin_of(cc, t.cblo[si.axis], t.cbhi[si.axis]) <= si.bin;
}

__global__ __launch_bounds__(256) void k_part_count(const BuildTask *tasks, const SplitInfo *split,
                                                    const uint32_t *chunk_task, const uint32_t *chunk_begin,
                                                    const uint32_t *idx, const float4 *cen, uint32_t *chunk_left,
                                                    const uint32_t *nchunks_p)
{
    __shared__ uint32_t cnt;
    const uint32_t nch = *nchunks_p;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        if (threadIdx.x == 0) cnt = 0;
        __syncthreads();
        const uint32_t  ti = chunk_task[c];
        const BuildTask t  = tasks[ti];
        const SplitInfo si = split[ti];
        const uint32_t  b  = chunk_begin[c];
        const uint32_t  e  = min(b + kChunk, t.end);
        uint32_t        local = 0;
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) local += goes_left(si, t, i, idx[i], cen) ? 1u : 0u;
        atomicAdd(&cnt, local);
        __syncthreads();
        if (threadIdx.x == 0) chunk_left[c] = cnt;
        __syncthreads();   // cnt is reset for the next chunk
    }
}

// Exclusive scan of the chunks' left counts within each task: one 256-thread block per task,
// 256 chunks at a time (a 14.8M-triangle root task has 7,200 chunks).
__global__ __launch_bounds__(256) void k_part_offsets(const uint32_t *ntasks_p, const uint32_t *task_chunk0,
                                                      const uint32_t *chunk_left, uint32_t *chunk_loff)
{
    __shared__ uint32_t wsum[4];
    const uint32_t ntasks = *ntasks_p;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const uint32_t c0 = task_chunk0[t], c1 = task_chunk0[t + 1];
        uint32_t carry = 0;
        for (uint32_t base = c0; base < c1; base += blockDim.x) {
            const uint32_t c = base + threadIdx.x;
            const uint32_t v = c < c1 ? chunk_left[c] : 0u;
            uint32_t incl = v;
            for (uint32_t o = 1; o < 64; o <<= 1) {
                const uint32_t x = __shfl_up(incl, o);
                if (lane >= o) incl += x;
            }
            if (lane == 63) wsum[w] = incl;
            __syncthreads();
            uint32_t before = 0, tot = 0;
            for (uint32_t j = 0; j < 4; ++j) {
                before += j < w ? wsum[j] : 0u;
                tot += wsum[j];
            }
            if (c < c1) chunk_loff[c] = carry + before + incl - v;
            carry += tot;
            __syncthreads();   // wsum is rewritten by the next tile
        }
    }
}

// Stable partition of one chunk: lefts in order, then rights in order.
__global__ __launch_bounds__(256) void k_part_scatter(const BuildTask *tasks, const SplitInfo *split,
                                                      const uint32_t *chunk_task, const uint32_t *chunk_begin,
                                                      const uint32_t *chunk_loff, const uint32_t *idx,
                                                      const float4 *cen, uint32_t *tmp, const uint32_t *nchunks_p)
{
    __shared__ uint32_t wsum[4];
    const uint32_t nch = *nchunks_p;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t  ti = chunk_task[c];
        const BuildTask t  = tasks[ti];
        const SplitInfo si = split[ti];
        const uint32_t  b  = chunk_begin[c];
        const uint32_t  e  = min(b + kChunk, t.end);
        const int       lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        uint32_t        lbefore = chunk_loff[c];            // lefts of this task before this chunk
        uint32_t        rbefore = (b - t.begin) - lbefore;  // rights before this chunk
        for (uint32_t base = b; base < e; base += blockDim.x) {
            const uint32_t i  = base + threadIdx.x;
            const bool     in = i < e;
            const uint32_t id = in ? idx[i] : 0u;
            const bool     f  = in && goes_left(si, t, i, id, cen);
            const uint64_t bal = __ballot(f);
            const uint32_t rank = __popcll(bal & ((1ull << lane) - 1ull));
            if (lane == 0) wsum[w] = __popcll(bal);
            __syncthreads();
            uint32_t off = 0, tot = 0;
            for (int j = 0; j < 4; ++j) {
                off += j < w ? wsum[j] : 0u;
                tot += wsum[j];
            }
            if (in) {
                const uint32_t lrank = off + rank;
                const uint32_t pos   = f ? t.begin + lbefore + lrank
                                         : t.begin + si.nleft + rbefore + (threadIdx.x - lrank);
                tmp[pos] = id;
            }
            const uint32_t m = min((uint32_t)blockDim.x, e - base);
            lbefore += tot;
            rbefore += m - tot;
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void k_copy_chunks(const BuildTask *tasks, const uint32_t *chunk_task,
                                                     const uint32_t *chunk_begin, const uint32_t *tmp, uint32_t *idx,
                                                     const uint32_t *nchunks_p)
{
    const uint32_t nch = *nchunks_p;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const BuildTask t = tasks[chunk_task[c]];
        const uint32_t  b = chunk_begin[c];
        const uint32_t  e = min(b + kChunk, t.end);
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) idx[i] = tmp[i];
    }
}

// ---- stage 1b: one workgroup builds a whole small segment -------------------------------
struct LocalTask {
    uint32_t b, e, slot, pad;
    float    cblo[3], cbhi[3];
};

constexpr uint32_t kWaveTask   = 128;   // subtrees of at most this many primitives: one wave each
constexpr uint32_t kMaxWTasks  = 512;   // wave-task list capacity (larger lists fall back to the block)
constexpr int      kWaveStack  = 8;     // per-wave stack: depth <= log2(kWaveTask)

struct WTask {
    uint32_t b, e, slot;
};

// Orders this wave's LDS accesses (lanes of one wave read what others wrote).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Subtrees of at most kSortMax primitives are split by an exact SAH sweep over their
// primitives sorted by centroid (one row of 16 lanes per axis) instead of binning: most binary
// nodes are this small, and the sweep costs a fraction of a binned step.
constexpr uint32_t kSortMax = 16;

// Position of lane pos's key among its row of 16 (ties by lane): count the smaller keys,
// visiting the row with DPP rotations (row_ror d: lane i reads lane (i - d) mod 16, which
// precedes i exactly when d <= i).
template <int d>
__device__ __forceinline__ uint32_t rank_in_row(float key, int pos, uint32_t r)
{
    const float kj = __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(
        (int)__float_as_uint(key), (int)__float_as_uint(key), 0x120 + d, 0xF, 0xF, false));
    r += (kj < key || (kj == key && d <= pos)) ? 1u : 0u;
    if constexpr (d < 15) return rank_in_row<d + 1>(key, pos, r);
    else return r;
}

// Inclusive prefix (suffix) box over a row of 16 lanes, Hillis-Steele with DPP row shifts;
// a lane whose source falls outside the row keeps its own value, which min / max ignore.
template <int d>
__device__ __forceinline__ void prefix_box(float *lo, float *hi)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], row_shift<0, d>(lo[k]));
        hi[k] = fmaxf(hi[k], row_shift<0, d>(hi[k]));
    }
    if constexpr (d < 8) prefix_box<2 * d>(lo, hi);
}
template <int d>
__device__ __forceinline__ void suffix_box(float *lo, float *hi)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], row_shift<1, d>(lo[k]));
        hi[k] = fmaxf(hi[k], row_shift<1, d>(hi[k]));
    }
    if constexpr (d < 8) suffix_box<2 * d>(lo, hi);
}

// One workgroup builds a whole small segment: nodes above kWaveTask primitives block-wide
// (one node at a time), then every wave takes whole subtrees of at most kWaveTask
// primitives from a shared list and builds them without block barriers.
__global__ __launch_bounds__(256) void k_small(const BuildTask *tasks, uint32_t ntasks, uint32_t *idx,
                                               const float4 *plo, const float4 *phi, const float4 *cen,
                                               Bvh2Node *nodes, uint32_t nslots, uint32_t *err, uint32_t max_leaf)
{
    __shared__ uint32_t  ids[kChunk];
    __shared__ uint32_t  tmp[kChunk];
    __shared__ int       bins[kBinWords];
    __shared__ SahOut    s;
    __shared__ LocalTask stack[48];
    __shared__ int       sp;
    __shared__ LocalTask cur;
    __shared__ uint32_t  wsum[4];
    __shared__ int       dec_leaf, dec_axis;
    __shared__ uint32_t  dec_nleft;
    __shared__ WTask     wtasks[kMaxWTasks];
    __shared__ uint32_t  wcount, wnext;
    __shared__ int       wbins[4][kBinWords];
    __shared__ SahOut    wsah[4];
    __shared__ LocalTask wstack[4][kWaveStack];

    const uint32_t ti = blockIdx.x;
    if (ti >= ntasks) return;
    const BuildTask task = tasks[ti];
    const uint32_t  gb   = task.begin;
    const uint32_t  n    = task.end - task.begin;
    if (task.end <= task.begin || n > kChunk) {
        if (threadIdx.x == 0) atomicOr(err, 1u);
        return;
    }
    const int       lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) ids[i] = idx[gb + i];
    if (threadIdx.x == 0) {
        wcount = 0;
        wnext  = 0;
        if (n <= kWaveTask) {
            wtasks[0] = WTask{0, n, task.slot};
            wcount    = 1;
            sp        = 0;
            cur.e     = 0xFFFFFFFFu;
        } else {
            LocalTask r;
            r.b    = 0;
            r.e    = n;
            r.slot = task.slot;
            r.pad  = 0;
            for (int k = 0; k < 3; ++k) {
                r.cblo[k] = task.cblo[k];
                r.cbhi[k] = task.cbhi[k];
            }
            stack[0] = r;
            sp       = 1;
            cur.e    = 0;
        }
    }
    __syncthreads();
    // ---- phase A: nodes above kWaveTask primitives, whole block per node ----
    for (;;) {
        if (threadIdx.x == 0 && sp > 0) cur = stack[--sp];
        __syncthreads();
        if (cur.e == 0xFFFFFFFFu) break;
        bins_init(bins, threadIdx.x, blockDim.x);
        __syncthreads();
        {
            const float4 cblo = lo4(cur.cblo), ks = scale4(cur.cblo, cur.cbhi);
            for (uint32_t i = cur.b + threadIdx.x; i < cur.e; i += blockDim.x) {
                const uint32_t id = ids[i];
                bins_add(bins, cen[id], plo[id], phi[id], cblo, ks);
            }
        }
        __syncthreads();
        if (w == 0) sah_wave(bins, &s);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t nl = 0;
            int      ax = -1;
            dec_leaf    = decide(s, max_leaf, nl, ax);
            dec_nleft   = nl;
            dec_axis    = ax;
            if (dec_leaf) write_leaf(nodes, nslots, err, cur.slot, s, gb + cur.b, cur.e - cur.b);
            else write_inner(nodes, nslots, err, cur.slot, s, nl);
        }
        __syncthreads();
        if (!dec_leaf && dec_axis >= 0) {
            uint32_t lcount = 0, rcount = 0;
            const int   ax  = dec_axis;
            const float alo = cur.cblo[ax], aks = bin_scale(alo, cur.cbhi[ax]);
            for (uint32_t base = cur.b; base < cur.e; base += blockDim.x) {
                const uint32_t i  = base + threadIdx.x;
                const bool     in = i < cur.e;
                bool           f  = false;
                uint32_t       id = 0;
                if (in) {
                    id = ids[i];
                    const float4 c  = cen[id];
                    const float  cc = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
                    f = bin_at(cc, alo, aks) <= s.bin;
                }
                const uint64_t bal  = __ballot(f);
                const uint32_t rank = __popcll(bal & ((1ull << lane) - 1ull));
                if (lane == 0) wsum[w] = __popcll(bal);
                __syncthreads();
                uint32_t off = 0, tot = 0;
                for (int j = 0; j < 4; ++j) {
                    off += j < w ? wsum[j] : 0u;
                    tot += wsum[j];
                }
                if (in) {
                    const uint32_t lr = off + rank;
                    tmp[f ? cur.b + lcount + lr : cur.b + dec_nleft + rcount + (threadIdx.x - lr)] = id;
                }
                const uint32_t m = min((uint32_t)blockDim.x, cur.e - base);
                lcount += tot;
                rcount += m - tot;
                __syncthreads();
            }
            for (uint32_t i = cur.b + threadIdx.x; i < cur.e; i += blockDim.x) ids[i] = tmp[i];
        }
        // every wave has finished reading cur before thread 0 replaces it below
        __syncthreads();
        if (threadIdx.x == 0) {
            if (!dec_leaf) {
                LocalTask l, r;
                l.b    = cur.b;
                l.e    = cur.b + dec_nleft;
                l.slot = cur.slot + 1;
                l.pad  = 0;
                r.b    = cur.b + dec_nleft;
                r.e    = cur.e;
                r.slot = cur.slot + 2 * dec_nleft;
                r.pad  = 0;
                for (int k = 0; k < 3; ++k) {
                    l.cblo[k] = dec_axis >= 0 ? s.lcblo[k] : cur.cblo[k];
                    l.cbhi[k] = dec_axis >= 0 ? s.lcbhi[k] : cur.cbhi[k];
                    r.cblo[k] = dec_axis >= 0 ? s.rcblo[k] : cur.cblo[k];
                    r.cbhi[k] = dec_axis >= 0 ? s.rcbhi[k] : cur.cbhi[k];
                }
                // larger child first so the smaller is processed next: stack depth <= log2(n)+1;
                // children of at most kWaveTask primitives go to the wave list while it has room
                auto put = [&](const LocalTask &ch) {
                    if (ch.e - ch.b <= kWaveTask && wcount < kMaxWTasks) wtasks[wcount++] = WTask{ch.b, ch.e, ch.slot};
                    else stack[sp++] = ch;
                };
                if (l.e - l.b >= r.e - r.b) {
                    put(l);
                    put(r);
                } else {
                    put(r);
                    put(l);
                }
            }
            if (sp == 0) cur.e = 0xFFFFFFFFu;
        }
        __syncthreads();
    }
    __syncthreads();
    // ---- phase B: subtrees of at most kWaveTask primitives, one wave each, no block barriers ----
    int       *wb  = wbins[w];
    SahOut    *ws  = &wsah[w];
    LocalTask *wst = wstack[w];
    const uint64_t lt = (1ull << lane) - 1ull;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&wnext, 1u);
        t = __shfl(t, 0);
        if (t >= wcount) break;
        const WTask wt = wtasks[t];
        if (wt.b >= wt.e || wt.e > n) {
            if (lane == 0) atomicOr(err, 1u);
            continue;
        }
        // centroid bounds of the subtree
        float clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
        for (uint32_t i = wt.b + lane; i < wt.e; i += 64) {
            const float4 c = cen[ids[i]];
            clo[0] = fminf(clo[0], c.x), clo[1] = fminf(clo[1], c.y), clo[2] = fminf(clo[2], c.z);
            chi[0] = fmaxf(chi[0], c.x), chi[1] = fmaxf(chi[1], c.y), chi[2] = fmaxf(chi[2], c.z);
        }
        for (int o = 32; o > 0; o >>= 1)
            for (int k = 0; k < 3; ++k) {
                clo[k] = fminf(clo[k], __shfl_xor(clo[k], o));
                chi[k] = fmaxf(chi[k], __shfl_xor(chi[k], o));
            }
        LocalTask c;
        c.b    = wt.b;
        c.e    = wt.e;
        c.slot = wt.slot;
        c.pad  = 0;
        for (int k = 0; k < 3; ++k) {
            c.cblo[k] = clo[k];
            c.cbhi[k] = chi[k];
        }
        int wsp = 0;   // wave-uniform
        for (;;) {
            const uint32_t cn     = c.e - c.b;
            const bool     sorted = cn <= kSortMax;
            int            leaf;
            uint32_t       nl = 0;
            int            ax = -1;
            bool           keep_l = true, keep_r = true;   // children still to be built
            if (cn == 2u) {
                // ---- two primitives: the only split is one against one.  Ordered by centroid
                // x, ties by position, with the cost summed in that order, it is what the
                // sweep below decides (its three rows tie, and row x has the lowest lanes) ----
                uint32_t id = 0;
                float    kx = 0.0f;
                float4   blo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), bhi = blo;
                if (lane < 2) {
                    id  = ids[c.b + lane];
                    kx  = cen[id].x;
                    blo = plo[id];
                    bhi = phi[id];
                }
                auto rl = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
                const bool     sw  = rl(kx, 1) < rl(kx, 0);    // the second comes first
                const int      f0  = sw ? 1 : 0, f1 = sw ? 0 : 1;
                const uint32_t id0 = (uint32_t)__builtin_amdgcn_readlane((int)id, 0);
                const uint32_t id1 = (uint32_t)__builtin_amdgcn_readlane((int)id, 1);
                SahOut a, b, so;
                const float la[3] = {rl(blo.x, 0), rl(blo.y, 0), rl(blo.z, 0)}, ha[3] = {rl(bhi.x, 0), rl(bhi.y, 0), rl(bhi.z, 0)};
                const float lb[3] = {rl(blo.x, 1), rl(blo.y, 1), rl(blo.z, 1)}, hb[3] = {rl(bhi.x, 1), rl(bhi.y, 1), rl(bhi.z, 1)};
                for (int k = 0; k < 3; ++k) {
                    a.lo[k]  = f0 ? lb[k] : la[k];
                    a.hi[k]  = f0 ? hb[k] : ha[k];
                    b.lo[k]  = f1 ? lb[k] : la[k];
                    b.hi[k]  = f1 ? hb[k] : ha[k];
                    so.lo[k] = fminf(a.lo[k], b.lo[k]);
                    so.hi[k] = fmaxf(a.hi[k], b.hi[k]);
                }
                so.n     = 2u;
                so.cost  = area3(a.lo[0], a.lo[1], a.lo[2], a.hi[0], a.hi[1], a.hi[2]) * 1.0f +
                           area3(b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]) * 1.0f;
                so.axis  = 0;
                so.bin   = 0;
                so.nleft = 1u;
                leaf     = decide(so, max_leaf, nl, ax);
                if (lane == 0) {
                    if (leaf) write_leaf(nodes, nslots, err, c.slot, so, gb + c.b, cn);
                    else write_inner(nodes, nslots, err, c.slot, so, nl);
                }
                if (!leaf && ax >= 0) {   // both children are one-primitive leaves
                    if (lane == 0) {
                        ids[c.b]     = sw ? id1 : id0;
                        ids[c.b + 1] = sw ? id0 : id1;
                        write_leaf(nodes, nslots, err, c.slot + 1u, a, gb + c.b, 1u);
                        write_leaf(nodes, nslots, err, c.slot + 2u, b, gb + c.b + 1u, 1u);
                    }
                    keep_l = keep_r = false;
                }
                wave_sync();
            } else if (sorted) {
                // ---- exact SAH sweep over the sorted primitives (at most 16) ----
                const int  row = lane >> 4, pos = lane & 15;   // row a: primitives sorted by axis a
                const bool val = lane < 48 && (uint32_t)pos < cn;
                uint32_t   id  = 0;
                float      key = kInf;
                float4     blo = make_float4(kInf, kInf, kInf, 0.0f), bhi = make_float4(-kInf, -kInf, -kInf, 0.0f);
                if (val) {
                    id = ids[c.b + pos];
                    const float4 cc = cen[id];
                    key = pick(cc.x, cc.y, cc.z, row);
                    blo = plo[id];
                    bhi = phi[id];
                }
                const uint32_t rank = rank_in_row<1>(key, pos, 0u);
                uint32_t *sid = reinterpret_cast<uint32_t *>(wb);        // [3][16] ids in sorted order
                float    *sbx = reinterpret_cast<float *>(wb + 48);      // [6][3][16] their boxes
                if (val) {
                    const int at = row * 16 + (int)rank;
                    sid[at]           = id;
                    sbx[0 * 48 + at]  = blo.x;
                    sbx[1 * 48 + at]  = blo.y;
                    sbx[2 * 48 + at]  = blo.z;
                    sbx[3 * 48 + at]  = bhi.x;
                    sbx[4 * 48 + at]  = bhi.y;
                    sbx[5 * 48 + at]  = bhi.z;
                }
                wave_sync();
                float pl[3], ph[3];   // box of sorted positions 0..pos
                for (int k = 0; k < 3; ++k) {
                    pl[k] = val ? sbx[k * 48 + lane] : kInf;
                    ph[k] = val ? sbx[(3 + k) * 48 + lane] : -kInf;
                }
                float sl[3] = {pl[0], pl[1], pl[2]}, sh[3] = {ph[0], ph[1], ph[2]};   // positions pos..15
                prefix_box<1>(pl, ph);
                suffix_box<1>(sl, sh);
                float rl[3], rh[3];   // box of positions pos + 1 .. cn - 1: right side of the split after pos
                for (int k = 0; k < 3; ++k) {
                    rl[k] = row_shift<1, 1>(sl[k]);
                    rh[k] = row_shift<1, 1>(sh[k]);
                }
                float cost = kInf;
                if (lane < 48 && (uint32_t)pos + 1u < cn)
                    cost = area3(pl[0], pl[1], pl[2], ph[0], ph[1], ph[2]) * (float)(pos + 1) +
                           area3(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]) * (float)(cn - 1u - (uint32_t)pos);
                // argmin over (cost, lane) as one unsigned key: the cost's bits (non-negative
                // floats order as integers) with the low 6 bits replaced by the lane, so costs
                // within 2^-17 of each other go to the lower lane
                uint32_t bk = (__float_as_uint(cost) & ~63u) | (uint32_t)lane;
                for (int o = 32; o > 0; o >>= 1) bk = min(bk, (uint32_t)__shfl_xor((int)bk, o));
                const int   bl = (int)(bk & 63u);
                const float bc = __shfl(cost, bl);
                SahOut so;
                const int last = (int)cn - 1;   // wave-uniform
                for (int k = 0; k < 3; ++k) {
                    so.lo[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pl[k]), last));
                    so.hi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ph[k]), last));
                }
                so.n     = cn;
                so.cost  = bc;
                so.axis  = bc == kInf ? -1 : bl >> 4;
                so.bin   = bc == kInf ? -1 : bl & 15;
                so.nleft = bc == kInf ? 0u : (uint32_t)(bl & 15) + 1u;
                leaf     = decide(so, max_leaf, nl, ax);
                if (lane == 0) {
                    if (leaf) write_leaf(nodes, nslots, err, c.slot, so, gb + c.b, cn);
                    else write_inner(nodes, nslots, err, c.slot, so, nl);
                }
                if (!leaf && ax >= 0) {
                    if (row == ax && (uint32_t)pos < cn) ids[c.b + pos] = sid[lane];
                    // a one-primitive child is a leaf: the splitting lane holds its box
                    keep_l = nl > 1u;
                    keep_r = cn - nl > 1u;
                    if (lane == bl) {
                        if (!keep_l) {
                            SahOut lb;
                            for (int k = 0; k < 3; ++k) lb.lo[k] = pl[k], lb.hi[k] = ph[k];
                            write_leaf(nodes, nslots, err, c.slot + 1u, lb, gb + c.b, 1u);
                        }
                        if (!keep_r) {
                            SahOut rb;
                            for (int k = 0; k < 3; ++k) rb.lo[k] = rl[k], rb.hi[k] = rh[k];
                            write_leaf(nodes, nslots, err, c.slot + 2u * nl, rb, gb + c.b + nl, 1u);
                        }
                    }
                }
                wave_sync();
            } else {
                bins_init(wb, lane, 64);
                wave_sync();
                const float4 cblo = lo4(c.cblo), ks = scale4(c.cblo, c.cbhi);
                for (uint32_t i = c.b + lane; i < c.e; i += 64) {
                    const uint32_t id = ids[i];
                    bins_add(wb, cen[id], plo[id], phi[id], cblo, ks);
                }
                wave_sync();
                sah_wave(wb, ws);
                wave_sync();
                leaf = decide(*ws, max_leaf, nl, ax);   // same result in every lane
                if (lane == 0) {
                    if (leaf) write_leaf(nodes, nslots, err, c.slot, *ws, gb + c.b, cn);
                    else write_inner(nodes, nslots, err, c.slot, *ws, nl);
                }
                if (!leaf && ax >= 0) {
                    const int      bin = ws->bin;
                    const float    alo = pick(c.cblo[0], c.cblo[1], c.cblo[2], ax);
                    const float    aks = bin_scale(alo, pick(c.cbhi[0], c.cbhi[1], c.cbhi[2], ax));
                    uint32_t       lcount = 0, rcount = 0;
                    for (uint32_t base = c.b; base < c.e; base += 64) {
                        const uint32_t i  = base + lane;
                        const bool     in = i < c.e;
                        bool           f  = false;
                        uint32_t       id = 0;
                        if (in) {
                            id = ids[i];
                            const float4 cc4 = cen[id];
                            const float  cc  = ax == 0 ? cc4.x : (ax == 1 ? cc4.y : cc4.z);
                            f = bin_at(cc, alo, aks) <= bin;
                        }
                        const uint64_t bal  = __ballot(f);
                        const uint32_t rank = __popcll(bal & lt);
                        if (in) tmp[f ? c.b + lcount + rank : c.b + nl + rcount + ((uint32_t)lane - rank)] = id;
                        const uint32_t tot = __popcll(bal);
                        lcount += tot;
                        rcount += min(64u, c.e - base) - tot;
                    }
                    wave_sync();
                    for (uint32_t i = c.b + lane; i < c.e; i += 64) ids[i] = tmp[i];
                    wave_sync();
                }
            }
            if (!leaf && (keep_l || keep_r)) {
                // children [b, b + nl) at slot + 1 and [b + nl, e) at slot + 2 nl.  With both
                // still to build, the larger waits and the smaller (at most half of c) is next:
                // every push halves the size, so at most log2(kWaveTask) = 7 entries are ever
                // held.  (Field-wise selects keep the tasks in registers.)  Children of a sorted
                // node are sorted nodes too and need no centroid bounds.
                const uint32_t lb = c.b, le = c.b + nl, rb = c.b + nl, re = c.e;
                const uint32_t lsl = c.slot + 1, rsl = c.slot + 2 * nl;
                const bool     lbig = keep_l && keep_r ? le - lb >= re - rb : keep_r;
                if (sorted) {   // wave-uniform: ranges and slots only
                    if (keep_l && keep_r) {
                        if (lane == 0) {
                            LocalTask &d = wst[wsp];
                            d.b    = lbig ? lb : rb;
                            d.e    = lbig ? le : re;
                            d.slot = lbig ? lsl : rsl;
                        }
                        ++wsp;
                    }
                    c.b    = lbig ? rb : lb;
                    c.e    = lbig ? re : le;
                    c.slot = lbig ? rsl : lsl;
                    wave_sync();
                    continue;
                }
                float lcl[3], lch[3], rcl[3], rch[3];
                const bool byax = ax >= 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    // values, not addresses, are selected (a selected address would put c in memory)
                    const float a0 = ws->lcblo[k], a1 = ws->lcbhi[k], a2 = ws->rcblo[k], a3 = ws->rcbhi[k];
                    const float c0 = c.cblo[k], c1 = c.cbhi[k];
                    lcl[k] = byax ? a0 : c0;
                    lch[k] = byax ? a1 : c1;
                    rcl[k] = byax ? a2 : c0;
                    rch[k] = byax ? a3 : c1;
                }
                if (keep_l && keep_r) {
                    if (lane == 0) {
                        LocalTask &d = wst[wsp];
                        d.b    = lbig ? lb : rb;
                        d.e    = lbig ? le : re;
                        d.slot = lbig ? lsl : rsl;
                        d.pad  = 0;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            d.cblo[k] = lbig ? lcl[k] : rcl[k];
                            d.cbhi[k] = lbig ? lch[k] : rch[k];
                        }
                    }
                    ++wsp;
                }
                // next: the smaller child, or the only one left to build (lbig = keep_r then)
                c.b    = lbig ? rb : lb;
                c.e    = lbig ? re : le;
                c.slot = lbig ? rsl : lsl;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c.cblo[k] = lbig ? rcl[k] : lcl[k];
                    c.cbhi[k] = lbig ? rch[k] : lch[k];
                }
                wave_sync();
                continue;
            }
            if (wsp == 0) break;
            --wsp;
            wave_sync();
            c = wst[wsp];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) idx[gb + i] = ids[i];
}

// ---- stage 2/3: collapse + encode ------------------------------------------------------
struct Frontier {
    uint32_t slot;   // BVH2 slot
    uint32_t wide;   // wide node index
};

__device__ __forceinline__ bool b2_leaf(const Bvh2Node &n) { return (n.b & kLeafBit) != 0u; }
__device__ __forceinline__ float b2_area(const Bvh2Node &n)
{
    return area3(n.lo[0], n.lo[1], n.lo[2], n.hi[0], n.hi[1], n.hi[2]);
}

// biased exponent e such that 255 * 2^(e-127) >= ext
__device__ __forceinline__ uint32_t quant_exp(double ext)
{
    if (!(ext > 0.0)) return 1u;
    int e;
    frexp(ext / 255.0, &e);   // ext/255 = m * 2^e, m in [0.5, 1) -> 2^e >= ext/255
    int be = e + 127;
    if (be < 1) be = 1;
    if (be > 254) be = 254;
    return (uint32_t)be;
}

// ---- SAH-optimal collapse of the binary tree into W-wide nodes (Ylitie, Karras and Laine
// 2017, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", sec. 3).
// C(n, i) is the least cost of representing the binary subtree of n by at most i child
// slots of a wide node: C(n, 1) = min(leaf, internal) with leaf = A(n) c_prim count(n) (when
// count(n) <= the leaf limit) and internal = A(n) c_node + min_j C(l, j) + C(r, W - j);
// C(n, i) = min(C(n, i - 1), min_j C(l, j) + C(r, i - j)).  The binary nodes are listed level
// by level top-down, then the costs are computed one level per launch bottom-up, so each
// launch reads what earlier launches wrote (no cross-workgroup synchronisation).
constexpr uint32_t kMaxLeafPrims = 16;   // a leaf reference holds count - 1 in 4 bits
// default cost of a primitive test, wide-node test = 1.  Measured against the greedy collapse
// (soup 1M / terrain 14.8M, 1080p Mrays/s: 2904 / 3093): 1.0 -> 2960 / 3163, 0.7 -> 2939 / 3175,
// 0.5 -> 2929 / 3173, 0.3 -> 2829 / 3248 (fewer node tests, more triangle tests per ray).
constexpr float    kPrimCost     = 1.0f;
constexpr uint32_t kMaxLevels    = 1024; // binary depth handled by the level lists

struct CollapseDP {
    float    *cost;     // [slot][8]: C(n, i) for i = 1..W
    uint32_t *choice;   // [slot]: 4 bits per i: i = 1: 0 leaf, else the internal split j;
                        //         i > 1: 0 = as for i - 1, else the split j (l gets j slots)
    uint32_t *count;    // [slot]: primitives under the node
    uint32_t *first;    // [slot]: position of its first primitive
};

// Level lists: list[off[L] .. off[L] + cnt[L]) holds the slots of depth L.  Launch L appends
// the children of its inner nodes as level L + 1 (cnt zeroed beforehand, cnt[0] = 1 with
// list[0] = the root, off[0] = 0).
constexpr uint32_t kLevelBlock = 256;

__global__ __launch_bounds__(kLevelBlock) void k_level_down(const Bvh2Node *nodes, uint32_t nslots, uint32_t *list,
                                                            uint32_t *cnt, uint32_t *off, uint32_t L)
{
    const uint32_t n = cnt[L], base = off[L], out = base + n;
    if (blockIdx.x == 0 && threadIdx.x == 0) off[L + 1] = out;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    __shared__ uint32_t wsum[kLevelBlock / 64], bfirst;
    // whole blocks step through the list together, so every block appends with one atomic
    // (a wave each made ~8K atomics on one counter per large level, serialised at the L2)
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        Bvh2Node nd{};
        bool inner = false;
        if (i < n) {
            nd    = nodes[list[base + i]];
            inner = !b2_leaf(nd) && nd.a < nslots && nd.b < nslots;
        }
        const uint64_t bal  = __ballot(inner);
        const uint32_t rank = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t j = 0; j < kLevelBlock / 64; ++j) {
            before += j < w ? wsum[j] : 0u;
            tot += wsum[j];
        }
        if (threadIdx.x == 0) bfirst = tot ? atomicAdd(cnt + L + 1, 2u * tot) : 0u;
        __syncthreads();
        const uint32_t k = bfirst + 2u * (before + rank);
        if (inner && out + k + 1 < nslots) {
            list[out + k]     = nd.a;
            list[out + k + 1] = nd.b;
        }
        __syncthreads();   // wsum and bfirst are rewritten by the next round
    }
}

template <int W>
__global__ void k_dp_level(const Bvh2Node *nodes, const uint32_t *list, const uint32_t *cnt, const uint32_t *off,
                           uint32_t L, CollapseDP dp, float c_node, float c_prim, uint32_t max_leaf)
{
    const uint32_t n = cnt[L], base = off[L];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s  = list[base + i];
        const Bvh2Node nd = nodes[s];
        const float    A  = b2_area(nd);
        if (b2_leaf(nd)) {
            const uint32_t c = nd.b & ~kLeafBit;
            const float    v = A * c_prim * (float)c;
            for (int k = 0; k < W; ++k) dp.cost[8 * (size_t)s + k] = v;
            dp.choice[s] = 0u;
            dp.count[s]  = c;
            dp.first[s]  = nd.a;
            continue;
        }
        float cl[W], cr[W];
        for (int k = 0; k < W; ++k) {
            cl[k] = dp.cost[8 * (size_t)nd.a + k];
            cr[k] = dp.cost[8 * (size_t)nd.b + k];
        }
        const uint32_t c = dp.count[nd.a] + dp.count[nd.b];
        float    best = __builtin_huge_valf();
        uint32_t bj   = 1;
        for (int j = 1; j < W; ++j) {
            const float v = cl[j - 1] + cr[W - j - 1];
            if (v < best) {
                best = v;
                bj   = (uint32_t)j;
            }
        }
        const float internal = A * c_node + best;
        const float leaf     = c <= max_leaf ? A * c_prim * (float)c : __builtin_huge_valf();
        float    C[W];
        uint32_t ch = leaf <= internal ? 0u : bj;
        C[0] = fminf(leaf, internal);
        for (int k = 2; k <= W; ++k) {
            float    b  = C[k - 2];
            uint32_t bc = 0;
            for (int j = 1; j < k; ++j) {
                const float v = cl[j - 1] + cr[k - j - 1];
                if (v < b) {
                    b  = v;
                    bc = (uint32_t)j;
                }
            }
            C[k - 1] = b;
            ch |= bc << (4 * (k - 1));
        }
        for (int k = 0; k < W; ++k) dp.cost[8 * (size_t)s + k] = C[k];
        dp.choice[s] = ch;
        dp.count[s]  = c;
        dp.first[s]  = dp.first[nd.a];
    }
}

template <int W>
// Called by every lane of a wave together (valid: the lane holds a frontier node); the wave
// allocates its nodes' inner children with one atomic per counter.
__device__ __forceinline__ void collapse_node(bool valid, const Bvh2Node *b2, const Frontier f, Frontier *out,
                                              uint32_t *nout, uint32_t *wide_count, void *wide_nodes,
                                              const CollapseDP &dp)
{
    // children: binary slots, each a leaf of its whole range (leafy) or a wide node to be
    uint32_t ch[W];
    bool     leafy[W];
    int      nc = 0;
    if (valid) {
        const Bvh2Node root = b2[f.slot];
        uint32_t c1 = dp.choice[f.slot] & 15u;
        // a binary leaf, or a root that the cost model makes one leaf (the root is the only
        // frontier node that can carry a leaf choice)
        const bool one_leaf = b2_leaf(root) || (c1 == 0u && dp.count[f.slot] <= kMaxLeafPrims);
        if (!one_leaf && (c1 == 0u || c1 >= (uint32_t)W)) c1 = (uint32_t)W / 2u;
        if (one_leaf) {
            ch[nc]      = f.slot;
            leafy[nc++] = true;
        } else {
            uint32_t sslot[W], sk[W];
            int      sp = 0;
            sslot[sp] = root.b;
            sk[sp++]  = (uint32_t)W - c1;
            sslot[sp] = root.a;
            sk[sp++]  = c1;
            while (sp > 0 && nc < W) {
                --sp;
                const uint32_t m = sslot[sp], k = sk[sp];
                const Bvh2Node mn = b2[m];
                if (b2_leaf(mn)) {
                    ch[nc]      = m;
                    leafy[nc++] = true;
                    continue;
                }
                const uint32_t cm = dp.choice[m];
                if (k == 1u) {
                    ch[nc]      = m;
                    leafy[nc++] = (cm & 15u) == 0u && dp.count[m] <= kMaxLeafPrims;
                    continue;
                }
                const uint32_t j = (cm >> (4u * (k - 1u))) & 15u;
                if (j == 0u || j >= k) {
                    sslot[sp] = m;
                    sk[sp++]  = k - 1u;
                } else if (sp + 2 <= W) {
                    sslot[sp] = mn.b;
                    sk[sp++]  = k - j;
                    sslot[sp] = mn.a;
                    sk[sp++]  = j;
                } else {   // cannot happen: the slots on the stack sum to at most W
                    ch[nc]      = m;
                    leafy[nc++] = false;
                }
            }
            while (sp > 0 && nc < W) {   // cannot happen either; keep every subtree in the tree
                --sp;
                ch[nc]      = sslot[sp];
                leafy[nc++] = b2_leaf(b2[sslot[sp]]);
            }
        }
    }
    uint32_t ref[W];
    float    lo[W][3], hi[W][3];
    bool     inner[W];
    uint32_t ninner = 0;
    for (int j = 0; j < nc; ++j) {
        const Bvh2Node n = b2[ch[j]];
        for (int k = 0; k < 3; ++k) {
            lo[j][k] = n.lo[k];
            hi[j][k] = n.hi[k];
        }
        inner[j] = !leafy[j];
        ninner += inner[j] ? 1u : 0u;
        if (!inner[j]) {
            const uint32_t cnt = b2_leaf(n) ? (n.b & ~kLeafBit) : dp.count[ch[j]];
            const uint32_t fst = b2_leaf(n) ? n.a : dp.first[ch[j]];
            ref[j] = ref_leaf(fst, cnt);
        }
    }
    // a node's inner children get consecutive indices, so siblings share DRAM pages and
    // translation entries; the wave's nodes take consecutive ranges (prefix sum of ninner,
    // at most W < 16, from one ballot per bit)
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt   = (1ull << lane) - 1ull;
    uint32_t       pre = 0, tot = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
        const uint64_t m = __ballot((ninner >> b) & 1u);
        pre += (uint32_t)__popcll(m & lt) << b;
        tot += (uint32_t)__popcll(m) << b;
    }
    uint32_t wbase = 0, fbase = 0;
    if (lane == 0 && tot) {
        wbase = atomicAdd(wide_count, tot);
        fbase = atomicAdd(nout, tot);
    }
    wbase = __shfl(wbase, 0);
    fbase = __shfl(fbase, 0);
    if (!valid) return;
    if (ninner) {
        uint32_t wi = wbase + pre, fo = fbase + pre;
        for (int j = 0; j < nc; ++j) {
            if (!inner[j]) continue;
            ref[j]    = wi;
            out[fo++] = Frontier{ch[j], wi++};
        }
    }
    if (W == 4) {
        Node4 nd;
        for (int j = 0; j < 4; ++j) {
            const bool v = j < nc;
            nd.lox[j]   = v ? lo[j][0] : 0.0f;
            nd.loy[j]   = v ? lo[j][1] : 0.0f;
            nd.loz[j]   = v ? lo[j][2] : 0.0f;
            nd.hix[j]   = v ? hi[j][0] : 0.0f;
            nd.hiy[j]   = v ? hi[j][1] : 0.0f;
            nd.hiz[j]   = v ? hi[j][2] : 0.0f;
            nd.child[j] = v ? ref[j] : kInvalidRef;
        }
        nd.nchild = (uint32_t)nc;
        nd.pad0 = nd.pad1 = nd.pad2 = 0;
        static_cast<Node4 *>(wide_nodes)[f.wide] = nd;
    } else {
        Node8Q nd;
        float  olo[3] = {kInf, kInf, kInf}, ohi[3] = {-kInf, -kInf, -kInf};
        for (int j = 0; j < nc; ++j)
            for (int k = 0; k < 3; ++k) {
                olo[k] = fminf(olo[k], lo[j][k]);
                ohi[k] = fmaxf(ohi[k], hi[j][k]);
            }
        uint32_t e[3];
        double   sc[3];
        for (int k = 0; k < 3; ++k) {
            e[k]  = quant_exp((double)ohi[k] - (double)olo[k]);
            sc[k] = ldexp(1.0, (int)e[k] - 127);
        }
        nd.ox   = olo[0];
        nd.oy   = olo[1];
        nd.oz   = olo[2];
        nd.meta = e[0] | (e[1] << 8) | (e[2] << 16) | ((uint32_t)nc << 24);
        uint8_t *qlo[3] = {nd.qlox, nd.qloy, nd.qloz};
        uint8_t *qhi[3] = {nd.qhix, nd.qhiy, nd.qhiz};
        for (int j = 0; j < 8; ++j) {
            nd.child[j] = j < nc ? ref[j] : kInvalidRef;
            for (int k = 0; k < 3; ++k) {
                if (j < nc) {
                    double ql = floor(((double)lo[j][k] - (double)olo[k]) / sc[k]);
                    double qh = ceil(((double)hi[j][k] - (double)olo[k]) / sc[k]);
                    ql = ql < 0.0 ? 0.0 : (ql > 255.0 ? 255.0 : ql);
                    qh = qh < 0.0 ? 0.0 : (qh > 255.0 ? 255.0 : qh);
                    qlo[k][j] = (uint8_t)ql;
                    qhi[k][j] = (uint8_t)qh;
                } else {
                    qlo[k][j] = 0;
                    qhi[k][j] = 0;
                }
            }
        }
        for (int j = 0; j < 8; ++j) nd.pad[j] = 0;
        static_cast<Node8Q *>(wide_nodes)[f.wide] = nd;
    }
}

// One wide level: the frontier's count is read from device memory (written by the level
// above), so levels are launched back to back without the host reading each count.
template <int W>
__global__ __launch_bounds__(256) void k_collapse(const Bvh2Node *b2, const Frontier *in, const uint32_t *nin_p,
                                                  Frontier *out, uint32_t *nout, uint32_t *wide_count,
                                                  void *wide_nodes, CollapseDP dp)
{
    const uint32_t nin = *nin_p;
    // the loop bound is the same for the whole wave, so its lanes call collapse_node together
    for (uint32_t w0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < nin; w0 += gridDim.x * blockDim.x) {
        const uint32_t i     = w0 + (threadIdx.x & 63u);
        const bool     valid = i < nin;
        collapse_node<W>(valid, b2, valid ? in[i] : Frontier{0u, 0u}, out, nout, wide_count, wide_nodes, dp);
    }
}

// Structural check of the finished wide BVH, so that a bad tree fails the build instead of
// the traversal: child counts in [1, W]; inner children after their parent (which also rules
// out cycles) and inside the pool; leaf ranges inside the triangle array; quantised child
// boxes with lo <= hi on every axis (the traversal picks entry planes by ray direction).
template <int W>
__global__ void k_check_wide(const void *wide, uint32_t nwide, uint32_t ntri, uint32_t *err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nwide) return;
    uint32_t nc;
    uint32_t ch[W];
    bool     ok = true;
    if constexpr (W == 4) {
        const Node4 &n = static_cast<const Node4 *>(wide)[i];
        nc = n.nchild;
        for (int j = 0; j < W; ++j) ch[j] = n.child[j];
    } else {
        const Node8Q &n = static_cast<const Node8Q *>(wide)[i];
        nc = n.meta >> 24;
        for (int j = 0; j < W; ++j) {
            ch[j] = n.child[j];
            if ((uint32_t)j < nc)
                ok = ok && n.qlox[j] <= n.qhix[j] && n.qloy[j] <= n.qhiy[j] && n.qloz[j] <= n.qhiz[j];
        }
    }
    ok = ok && nc >= 1u && nc <= (uint32_t)W;
    for (int j = 0; j < W; ++j) {
        if ((uint32_t)j >= nc) break;
        const uint32_t r = ch[j];
        if (ref_is_leaf(r)) {
            const uint32_t first = leaf_first(r), cnt = leaf_count(r);
            ok = ok && first < ntri && cnt <= ntri - first;
        } else {
            ok = ok && r > i && r < nwide;
        }
    }
    if (!ok) atomicOr(err, 1u);
}

__global__ void k_reorder(const uint32_t *idx, uint32_t n, const float *tri, float4 *out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t id = idx[p];
    const float   *q  = tri + 9 * (size_t)id;
    out[3 * (size_t)p + 0] = make_float4(q[0], q[1], q[2], __builtin_bit_cast(float, id));
    out[3 * (size_t)p + 1] = make_float4(q[3], q[4], q[5], 0.0f);
    out[3 * (size_t)p + 2] = make_float4(q[6], q[7], q[8], 0.0f);
}

__global__ void k_mark_unused(Bvh2Node *nodes, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) nodes[i].b = 0xFFFFFFFEu;
}

// SAH cost of the binary tree: sum(inner area) + sum(leaf area * count), / root area
__global__ void k_sah_cost(const Bvh2Node *nodes, uint32_t nslots, double *acc, uint32_t *count)
{
    double   local = 0.0;
    uint32_t used  = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) {
        const Bvh2Node n = nodes[i];
        if (n.b == 0xFFFFFFFEu) continue;
        ++used;
        const double a = (double)b2_area(n);
        local += b2_leaf(n) ? a * (double)(n.b & ~kLeafBit) : a;
    }
    atomicAdd(acc, local);
    atomicAdd(count, used);
}

#define HIP_TRY(x)                          \
    do {                                    \
        hipError_t e_ = (x);                \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

// The build's scratch memory: one device buffer per device, kept between builds and grown
// when a build needs more, up to kKeepBytes (≈ 4M triangles); a larger build allocates and
// frees its own.  Allocating the scratch per build cost ≈ 0.5 ms of host time in 31 pool
// allocations and ~0.25 ms of idle GPU when made while kernels were queued; with one pool
// allocation per build, some builds still spent 35 ms (once ~1 s) before their first kernel.
// Builds on one device take turns on the buffer.
constexpr uint64_t kKeepBytes  = 2ull << 30;
constexpr int      kMaxDevices = 64;

struct ScratchArena {
    std::mutex mu;
    void      *base  = nullptr;
    size_t     bytes = 0;
};

ScratchArena &device_arena(int dev)
{
    static ScratchArena arenas[kMaxDevices];   // device memory held until the process ends
    return arenas[dev];
}

// The build's scratch buffers, carved out of one allocation in 256-byte aligned pieces.
struct Arena {
    std::vector<std::function<void(char *)>> pieces;   // each sets its pointer from the base
    size_t                                   bytes = 0;
    template <class T>
    void add(T **p, size_t n)
    {
        const size_t off = bytes;
        pieces.push_back([p, off](char *base) { *p = static_cast<T *>(static_cast<void *>(base + off)); });
        bytes += (n + 255) & ~(size_t)255;
    }
    void place(void *base)
    {
        for (auto &f : pieces) f(static_cast<char *>(base));
    }
};

}  // namespace

hipError_t build_bvh(const float *d_tri, uint32_t ntri, uint32_t width, uint32_t max_leaf, hipStream_t s,
                     BuildResult *out, float prim_cost)
{
    *out = BuildResult{};
    if (ntri == 0) return hipSuccess;
    if (width != 4 && width != 8) return hipErrorInvalidValue;
    if (ntri >= (1u << 27)) return hipErrorInvalidValue;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 16) max_leaf = 16;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    ScratchArena               &sa = device_arena(dev);
    std::lock_guard<std::mutex> arena_lock(sa.mu);
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventRecord(ev0, s));

    const uint32_t nslots = 2 * ntri;
    float4 *plo, *phi, *cen;
    uint32_t *idx, *tmp;
    int *root;
    Bvh2Node *nodes;
    BuildTask *large_a, *large_b, *small;
    uint32_t *counts;   // [1] small tasks, [4] wide count, [5] error flag, [8..24) per-level rings
    const uint32_t max_large = ntri / kChunk + 2;
    const uint32_t max_chunks = ntri / kChunk + 2 * max_large + 2;
    uint32_t *chunk_task, *chunk_begin, *task_chunk0, *chunk_left, *chunk_loff;
    SplitInfo *split;
    int *gbins;
    Arena A;
    A.add(&plo, sizeof(float4) * ntri);
    A.add(&phi, sizeof(float4) * ntri);
    A.add(&cen, sizeof(float4) * ntri);
    A.add(&idx, sizeof(uint32_t) * ntri);
    A.add(&tmp, sizeof(uint32_t) * ntri);
    A.add(&root, sizeof(int) * 12);
    A.add(&nodes, sizeof(Bvh2Node) * nslots);
    A.add(&large_a, sizeof(BuildTask) * max_large);
    A.add(&large_b, sizeof(BuildTask) * max_large);
    A.add(&small, sizeof(BuildTask) * (ntri + 1));
    A.add(&counts, sizeof(uint32_t) * 32);
    A.add(&chunk_task, sizeof(uint32_t) * max_chunks);
    A.add(&chunk_begin, sizeof(uint32_t) * max_chunks);
    A.add(&task_chunk0, sizeof(uint32_t) * (max_large + 1));
    A.add(&chunk_left, sizeof(uint32_t) * max_chunks);
    A.add(&chunk_loff, sizeof(uint32_t) * max_chunks);
    A.add(&split, sizeof(SplitInfo) * max_large);
    A.add(&gbins, sizeof(int) * kBinWords * max_large);
    // scratch of the later stages
    double   *sah_acc;
    uint32_t *used;
    A.add(&sah_acc, sizeof(double));
    A.add(&used, sizeof(uint32_t));
    CollapseDP dp{};
    uint32_t  *lvl_list = nullptr, *lvl_cnt = nullptr, *lvl_off = nullptr;
    A.add(&dp.cost, sizeof(float) * 8 * (size_t)nslots);
    A.add(&dp.choice, sizeof(uint32_t) * (size_t)nslots);
    A.add(&dp.count, sizeof(uint32_t) * (size_t)nslots);
    A.add(&dp.first, sizeof(uint32_t) * (size_t)nslots);
    A.add(&lvl_list, sizeof(uint32_t) * (size_t)nslots);
    A.add(&lvl_cnt, sizeof(uint32_t) * (kMaxLevels + 1));
    A.add(&lvl_off, sizeof(uint32_t) * (kMaxLevels + 1));
    const uint32_t max_wide   = ntri + 1;
    const size_t   node_bytes = 128;
    void          *wide_tmp;
    Frontier      *fa, *fb;
    uint32_t      *fcnt;
    A.add(&wide_tmp, node_bytes * max_wide);
    A.add(&fa, sizeof(Frontier) * max_wide);
    A.add(&fb, sizeof(Frontier) * max_wide);
    A.add(&fcnt, sizeof(uint32_t) * ((size_t)max_wide + 6));
    void *arena     = nullptr;
    bool  own_arena = A.bytes > kKeepBytes;   // too large to keep: this build's own
    if (own_arena) {
        HIP_TRY(hipMalloc(&arena, A.bytes));
    } else {
        if (sa.bytes < A.bytes) {   // the previous build on this device has completed
            if (sa.base) HIP_TRY(hipFree(sa.base));
            sa.base  = nullptr;
            sa.bytes = 0;
            const size_t want = std::min<size_t>(kKeepBytes, A.bytes + A.bytes / 4);
            HIP_TRY(hipMalloc(&sa.base, want));
            sa.bytes = want;
        }
        arena = sa.base;
    }
    A.place(arena);
    float4 *tri_out = nullptr;   // the triangles in leaf order: part of the result

    // root bounds start empty (order-preserving integer images of +inf / -inf); small
    // constants are set with memsets, not copied from pageable host memory, which would wait
    // for the stream to drain
    const int o_inf = f2ord(kInf), o_ninf = f2ord(-kInf);
    for (int k = 0; k < 4; ++k) HIP_TRY(hipMemsetD32Async(root + 3 * k, (k & 1) ? o_ninf : o_inf, 3, s));
    HIP_TRY(hipMemsetAsync(counts, 0, sizeof(uint32_t) * 32, s));
    k_mark_unused<<<(nslots + 255) / 256, 256, 0, s>>>(nodes, nslots);
    const uint32_t nb = std::min<uint32_t>((ntri + 255) / 256, 2048u);
    k_prim_bounds<<<nb, 256, 0, s>>>(d_tri, ntri, plo, phi, cen, idx, root);
    // Per-level counters in rings of kRing entries: tcnt[L % kRing] large tasks and
    // ccnt[L % kRing] chunks of level L.  Levels are launched kLevelsPerRead at a time with
    // fixed grids; the host then reads the rings once and stops when the next level has no
    // large task (a level launched past the last one finds zero tasks and does nothing).
    constexpr uint32_t kLevelsPerRead = 4, kRing = 8;
    uint32_t          *tcnt = counts + 8, *ccnt = counts + 8 + kRing;
    k_root_task<<<1, 1, 0, s>>>(root, ntri, large_a, small, tcnt, counts);
    k_make_chunks<<<1, kChunkBlock, 0, s>>>(large_a, tcnt, chunk_task, chunk_begin, task_chunk0, ccnt);
    HIP_TRY(hipGetLastError());
    uint32_t       hc[8 + 2 * kRing];
    const uint32_t gt = std::min<uint32_t>(max_large, 1024u), gc = std::min<uint32_t>(max_chunks, 2048u);
    uint32_t L = 0;
    for (;;) {
        for (const uint32_t upto = L + kLevelsPerRead; L < upto; ++L) {
            BuildTask     *cur = (L & 1u) ? large_b : large_a, *nxt = (L & 1u) ? large_a : large_b;
            const uint32_t *nt = tcnt + L % kRing, *nc = ccnt + L % kRing;
            uint32_t       *nt1 = tcnt + (L + 1) % kRing, *nc1 = ccnt + (L + 1) % kRing;
            k_clear_bins<<<gt, 256, 0, s>>>(gbins, nt, nt1);
            k_bin_large<<<gc, 256, 0, s>>>(cur, chunk_task, chunk_begin, idx, plo, phi, cen, gbins, nc);
            k_split_large<<<gt, 64, 0, s>>>(cur, nt, gbins, split, nodes, nslots, nxt, small, nt1, counts, max_leaf);
            k_part_count<<<gc, 256, 0, s>>>(cur, split, chunk_task, chunk_begin, idx, cen, chunk_left, nc);
            k_part_offsets<<<gt, 256, 0, s>>>(nt, task_chunk0, chunk_left, chunk_loff);
            k_part_scatter<<<gc, 256, 0, s>>>(cur, split, chunk_task, chunk_begin, chunk_loff, idx, cen, tmp, nc);
            k_copy_chunks<<<gc, 256, 0, s>>>(cur, chunk_task, chunk_begin, tmp, idx, nc);
            // chunks of the next level (the chunk lists above have been consumed)
            k_make_chunks<<<1, kChunkBlock, 0, s>>>(nxt, nt1, chunk_task, chunk_begin, task_chunk0, nc1);
        }
        HIP_TRY(hipGetLastError());
        // the result's triangle buffer is allocated while the first levels run
        if (!tri_out) HIP_TRY(hipMalloc((void **)&tri_out, sizeof(float4) * 3 * (size_t)ntri));
        HIP_TRY(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (hc[8 + L % kRing] == 0u) break;
    }
    const uint32_t large_levels = L;   // an upper bound: the last batch may end in empty levels
    const uint32_t nsmall       = hc[1];
    if (nsmall) k_small<<<nsmall, 256, 0, s>>>(small, nsmall, idx, plo, phi, cen, nodes, nslots, counts + 5, max_leaf);
    HIP_TRY(hipGetLastError());
    // The SAH cost, the collapse's initialisation and the first batch of level lists are
    // queued behind k_small without waiting for it: they only read node slots inside the tree
    // (k_level_down checks every child slot), and k_small's error flag is read with the
    // level counts, before anything that trusts the tree runs.

    // SAH cost + used slots
    HIP_TRY(hipMemsetAsync(sah_acc, 0, sizeof(double), s));
    HIP_TRY(hipMemsetAsync(used, 0, sizeof(uint32_t), s));
    k_sah_cost<<<std::min<uint32_t>((nslots + 255) / 256, 1024u), 256, 0, s>>>(nodes, nslots, sah_acc, used);

    // SAH-optimal distribution of the binary tree over wide nodes (cost of a primitive test
    // relative to a wide-node test): level lists top-down, costs bottom-up
    // nodes the level lists do not reach (trees deeper than kMaxLevels) keep a split into
    // halves and are never turned into leaves
    HIP_TRY(hipMemsetAsync(dp.choice, 0x44, sizeof(uint32_t) * (size_t)nslots, s));
    HIP_TRY(hipMemsetAsync(dp.count, 0xFF, sizeof(uint32_t) * (size_t)nslots, s));
    HIP_TRY(hipMemsetAsync(lvl_cnt, 0, sizeof(uint32_t) * (kMaxLevels + 1), s));
    HIP_TRY(hipMemsetAsync(lvl_off, 0, sizeof(uint32_t) * (kMaxLevels + 1), s));
    HIP_TRY(hipMemsetAsync(lvl_list, 0, sizeof(uint32_t), s));   // level 0: the root, slot 0
    HIP_TRY(hipMemsetD32Async(lvl_cnt, 1, 1, s));
    uint32_t nlevels = 0;
    {
        // level L holds at most 2^L nodes: the grid of a level's launch is sized to that
        const uint32_t g    = std::min<uint32_t>((nslots + 255) / 256, 1024u);
        auto           grid = [&](uint32_t L) { return L >= 18u ? g : std::min(g, ((1u << L) + 255u) / 256u); };
        std::vector<uint32_t> lc(kMaxLevels + 1);
        // first batch: the large levels plus the depth a 2048-primitive segment usually reaches
        for (uint32_t L = 0, batch = large_levels + 24u; L < kMaxLevels && nlevels == 0; batch = 16u) {
            const bool     first = L == 0;
            const uint32_t upto  = std::min<uint32_t>(L + batch, kMaxLevels);
            for (; L < upto; ++L) k_level_down<<<grid(L), kLevelBlock, 0, s>>>(nodes, nslots, lvl_list, lvl_cnt, lvl_off, L);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(lc.data(), lvl_cnt, sizeof(uint32_t) * (L + 1), hipMemcpyDeviceToHost, s));
            if (first) HIP_TRY(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (first && hc[5] != 0u) {
                // a node slot or child range fell outside the tree: nothing past this point is safe
                HIP_TRY(hipStreamSynchronize(s));
                if (own_arena) hipFree(arena);
                hipFree(tri_out);
                hipEventDestroy(ev0);
                hipEventDestroy(ev1);
                out->invalid = 1;
                return hipErrorUnknown;
            }
            for (uint32_t k = 0; k <= L; ++k)
                if (lc[k] == 0u) {
                    nlevels = k;
                    break;
                }
        }
        if (nlevels == 0) nlevels = kMaxLevels;   // deeper trees keep a wide node per binary level below
        const float    cp   = prim_cost > 0.0f ? prim_cost : kPrimCost;
        const uint32_t lmax = std::min<uint32_t>(std::max<uint32_t>(max_leaf, 1u), kMaxLeafPrims);
        for (uint32_t L = nlevels; L-- > 0;) {
            if (width == 4)
                k_dp_level<4><<<grid(L), 256, 0, s>>>(nodes, lvl_list, lvl_cnt, lvl_off, L, dp, 1.0f, cp, lmax);
            else k_dp_level<8><<<grid(L), 256, 0, s>>>(nodes, lvl_list, lvl_cnt, lvl_off, L, dp, 1.0f, cp, lmax);
        }
        HIP_TRY(hipGetLastError());
    }

    // collapse
    HIP_TRY(hipMemsetAsync(fa, 0, sizeof(Frontier), s));   // the root: binary slot 0, wide node 0
    HIP_TRY(hipMemsetD32Async(counts + 4, 1, 1, s));
    // frontier sizes per wide level: fcnt[d] nodes at depth d (fcnt[0] = 1, the root).  The
    // depth is at most the number of wide nodes, so fcnt never overflows.  Levels are
    // launched in batches and the host reads the counts once per batch; a launch past the
    // last level finds an empty frontier and does nothing.
    HIP_TRY(hipMemsetD32Async(fcnt, 1, 1, s));
    uint32_t depth = 0;
    {
        const uint32_t gcap = std::min<uint32_t>((max_wide + 255) / 256, 1024u);
        auto cgrid = [&](uint32_t d) {   // depth d holds at most width^d nodes
            uint64_t b = 1;
            for (uint32_t k = 0; k < d && b < 256ull * gcap; ++k) b *= width;
            return (uint32_t)std::min<uint64_t>(gcap, (b + 255) / 256);
        };
        std::vector<uint32_t> fc;
        // first batch: about the wide depth that the binary depth implies (a wide level
        // spans ~log2(width) binary levels), then four levels at a time
        for (uint32_t d = 0, batch = nlevels / (width == 8 ? 3u : 2u) + 2u; depth == 0 && d < max_wide; batch = 4u) {
            const uint32_t upto = std::min<uint32_t>(d + batch, max_wide);
            HIP_TRY(hipMemsetAsync(fcnt + d + 1, 0, sizeof(uint32_t) * (upto - d), s));
            for (; d < upto; ++d) {
                const Frontier *in = (d & 1u) ? fb : fa;
                Frontier       *ou = (d & 1u) ? fa : fb;
                if (width == 4)
                    k_collapse<4><<<cgrid(d), 256, 0, s>>>(nodes, in, fcnt + d, ou, fcnt + d + 1, counts + 4, wide_tmp, dp);
                else
                    k_collapse<8><<<cgrid(d), 256, 0, s>>>(nodes, in, fcnt + d, ou, fcnt + d + 1, counts + 4, wide_tmp, dp);
            }
            HIP_TRY(hipGetLastError());
            fc.resize(d + 1);
            HIP_TRY(hipMemcpyAsync(fc.data(), fcnt, sizeof(uint32_t) * (d + 1), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (uint32_t k = 0; k <= d; ++k)
                if (fc[k] == 0u) {
                    depth = k;
                    break;
                }
        }
    }
    const uint32_t nwide = hc[4];
    if (width == 4) k_check_wide<4><<<(nwide + 255) / 256, 256, 0, s>>>(wide_tmp, nwide, ntri, counts + 5);
    else k_check_wide<8><<<(nwide + 255) / 256, 256, 0, s>>>(wide_tmp, nwide, ntri, counts + 5);
    k_reorder<<<(ntri + 255) / 256, 256, 0, s>>>(idx, ntri, d_tri, tri_out);
    HIP_TRY(hipGetLastError());
    void *wide;   // allocated while the two kernels above run
    HIP_TRY(hipMalloc(&wide, node_bytes * nwide));
    HIP_TRY(hipMemcpyAsync(wide, wide_tmp, node_bytes * nwide, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipEventRecord(ev1, s));

    double   sah_h = 0.0;
    uint32_t used_h = 0;
    HIP_TRY(hipMemcpyAsync(&sah_h, sah_acc, sizeof sah_h, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&used_h, used, sizeof used_h, hipMemcpyDeviceToHost, s));
    Bvh2Node root_node;
    HIP_TRY(hipMemcpyAsync(&root_node, nodes, sizeof root_node, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const bool invalid = hc[5] != 0u;   // the structure check of the wide tree failed
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
    const float rdx = root_node.hi[0] - root_node.lo[0], rdy = root_node.hi[1] - root_node.lo[1],
                rdz = root_node.hi[2] - root_node.lo[2];
    const double ra = 2.0 * ((double)rdx * rdy + (double)rdy * rdz + (double)rdz * rdx);

    if (own_arena) HIP_TRY(hipFree(arena));
    if (invalid) {
        hipFree(wide);
        hipFree(tri_out);
        out->invalid = 1;
        return hipErrorUnknown;
    }

    out->nodes      = wide;
    out->tri        = tri_out;
    out->node_count = nwide;
    out->bvh2_nodes = used_h;
    out->depth      = depth;
    out->root       = 0;
    out->sah        = ra > 0.0 ? sah_h / ra : 0.0;
    out->ms         = ms;
    return hipSuccess;
}

namespace {

template <int W>
__global__ void k_rebase(void *nodes, uint32_t count, uint32_t node_base, uint32_t tri_base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t *child;
    uint32_t  nc;
    if constexpr (W == 4) {
        Node4 *n = static_cast<Node4 *>(nodes) + i;
        child    = n->child;
        nc       = n->nchild;
    } else {
        Node8Q *n = static_cast<Node8Q *>(nodes) + i;
        child     = n->child;
        nc        = n->meta >> 24;
    }
    for (uint32_t j = 0; j < nc && j < (uint32_t)W; ++j) {
        const uint32_t r = child[j];
        child[j] = ref_is_leaf(r) ? ref_leaf(leaf_first(r) + tri_base, leaf_count(r)) : r + node_base;
    }
}

__global__ void k_add_prim_base(float4 *tri, uint32_t ntri, uint32_t base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri) return;
    float4 a = tri[3 * (size_t)i];
    a.w      = __uint_as_float(__float_as_uint(a.w) + base);
    tri[3 * (size_t)i] = a;
}

}  // namespace

hipError_t bvh_rebase(void *nodes, uint32_t count, uint32_t width, uint32_t node_base, uint32_t tri_base,
                      hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const uint32_t b = (count + 255) / 256;
    if (width == 4) k_rebase<4><<<b, 256, 0, s>>>(nodes, count, node_base, tri_base);
    else k_rebase<8><<<b, 256, 0, s>>>(nodes, count, node_base, tri_base);
    return hipGetLastError();
}

hipError_t tri_add_prim_base(float4 *tri, uint32_t ntri, uint32_t base, hipStream_t s)
{
    if (ntri == 0 || base == 0) return hipSuccess;
    k_add_prim_base<<<(ntri + 255) / 256, 256, 0, s>>>(tri, ntri, base);
    return hipGetLastError();
}

}  // namespace crt

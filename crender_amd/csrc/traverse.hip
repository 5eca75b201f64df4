// traverse.hip -- wide-BVH ray traversal for the crt path tracer (gfx950).
//
// Replaces rtcIntersect1 (CRender src/objects/model.cpp:27): closest hit in (tnear, tfar]
// with equal t resolved to the smaller primID; plus an any-hit variant for shadow rays.
//
// Execution model (wave64):
//   * persistent grid: a fixed number of 256-thread blocks per CU; every wave loops
//     fetching rays until the queue is drained, so the launch shape never depends on the
//     ray count (which lives in device memory and is produced by the previous kernel).
//   * XCD-aware fetch: the ray range is cut into 8 contiguous slices, one per XCD
//     (blocks are dispatched round-robin over XCDs), so neighbouring rays share an L2;
//     an XCD that drains its slice steals from the others.
//   * refill: when 32 or more lanes of a wave are idle, one lane reserves a run of rays
//     for all idle lanes with a single atomic; lanes take their ray by ballot rank.
//   * per-lane short stack in LDS (entry-major (ref, distance) pairs; 8, 16 or 32 entries), spilled
//     a full stack at a time to a per-lane global region when full.  Entries carry the child's entry
//     distance so popped nodes beyond the current hit are skipped.
//   * children are ordered by entry distance with a sorting network; the nearest is
//     visited next and the others pushed far-to-near, with unconditional LDS writes.
//   * leaf parking: a lane that reaches a leaf keeps it aside (up to two) and continues
//     down the tree until most lanes of the wave hold their leaves (fewer than 16 can still
//     park one), so the triangle phase runs with more lanes busy (Aila and Laine 2009,
//     "speculative traversal").
#include <hip/hip_runtime.h>
#include "crt_internal.h"

namespace crt {

namespace {

constexpr float kInf = __builtin_huge_valf();
constexpr uint32_t kEmpty = kInvalidRef;
constexpr uint32_t kMissKey = 0xFFFFFFFFu;
constexpr uint32_t kRefillIdle = 32;           // default idle lanes that trigger a refill
// default: a wave leaves the inner-node phase for its parked leaves once fewer than 16 lanes
// can still park one.  Measured on the 1M soup / 14.8M terrain at 1080p with two wavefronts in
// flight: 4 -> 2690 / 3031 Mrays/s, 8 -> 2814 / 3090, 16 -> 2852-2861 / 3079, 32 -> 2758 / 2979.
constexpr uint32_t kParkMin    = 16;
// default: after 3 leaf-loop rounds, once every parked leaf is tested, a wave resumes the
// inner-node phase (lanes still at a leaf park it there) instead of following chains of
// popped leaves with few lanes busy.  Soup / terrain: unlimited 2870 / 3048 Mrays/s, 1 round
// 2880 / 3055, 2 rounds 2891 / 3105, 3 rounds 2904 / 3093.
constexpr uint32_t kLeafRounds = 3;
constexpr uint32_t kInstExit   = 0xFFFFFFFEu;   // stack marker: leave the current instance
// Ize, "Robust BVH Ray Traversal" (JCGT 2013): scaling the far slab distance by
// 1 + 2*gamma(3) keeps FP32 box tests conservative.
constexpr float kFarScale = 1.0000003576278687f;

struct RayState {
    f3     o, d, inv;
    float  tn, tmax;
    wt_ray wt;
    uint32_t best_tri, best_prim;
    float  bu, bv;
};

// compare-exchange of (key, ref) pairs; refs travel with their keys so that every array
// index stays static and the arrays live in registers
__device__ __forceinline__ void cswap(uint32_t &a, uint32_t &b, uint32_t &ra, uint32_t &rb)
{
    const bool     sw = b < a;
    const uint32_t ka = sw ? b : a, kb = sw ? a : b;
    const uint32_t xa = sw ? rb : ra, xb = sw ? ra : rb;
    a  = ka;
    b  = kb;
    ra = xa;
    rb = xb;
}

template <int W>
__device__ __forceinline__ void sort_pairs(uint32_t (&k)[W], uint32_t (&r)[W])
{
#define CS(i, j) cswap(k[i], k[j], r[i], r[j])
    if constexpr (W == 4) {
        CS(0, 1); CS(2, 3);
        CS(0, 2); CS(1, 3);
        CS(1, 2);
    } else {
        // optimal 19-comparator network for 8 inputs
        CS(0, 2); CS(1, 3); CS(4, 6); CS(5, 7);
        CS(0, 4); CS(1, 5); CS(2, 6); CS(3, 7);
        CS(0, 1); CS(2, 3); CS(4, 5); CS(6, 7);
        CS(2, 4); CS(3, 5);
        CS(1, 4); CS(3, 6);
        CS(1, 2); CS(3, 4); CS(5, 6);
    }
#undef CS
}

// key = entry distance (non-negative float bits, low 3 bits cleared) | child slot
__device__ __forceinline__ uint32_t make_key(float t, int j)
{
    return (__float_as_uint(t) & ~7u) | (uint32_t)j;
}

// Finite reciprocal: a zero direction component gives +-1e30 instead of inf, so an
// axis-parallel ray whose origin lies on a slab plane yields [0, big] rather than the
// NaN of 0 * inf, and the box test stays conservative.
__device__ __forceinline__ float safe_rcp(float d)
{
    return fabsf(d) < 1e-30f ? copysignf(1e30f, d) : 1.0f / d;
}

__device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz, const RayState &r,
                                     float &tentry)
{
    float t0x = (lx - r.o.x) * r.inv.x, t1x = (hx - r.o.x) * r.inv.x;
    float t0y = (ly - r.o.y) * r.inv.y, t1y = (hy - r.o.y) * r.inv.y;
    float t0z = (lz - r.o.z) * r.inv.z, t1z = (hz - r.o.z) * r.inv.z;
    const float tnx = fminf(t0x, t1x), tfx = fmaxf(t0x, t1x);
    const float tny = fminf(t0y, t1y), tfy = fmaxf(t0y, t1y);
    const float tnz = fminf(t0z, t1z), tfz = fmaxf(t0z, t1z);
    const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, fmaxf(r.tn, 0.0f)));
    const float tmax = fminf(fminf(tfx, tfy), tfz) * kFarScale;
    tentry = tmin;
    return tmin <= fminf(tmax, r.tmax);
}

template <int W>
struct NodeOps;

template <>
struct NodeOps<4> {
    static __device__ __forceinline__ int expand(const void *nodes, uint32_t idx, const RayState &r, uint32_t (&key)[4],
                                                 uint32_t (&ref)[4])
    {
        const float4 *p = reinterpret_cast<const float4 *>(static_cast<const Node4 *>(nodes) + idx);
        const float4  lx = p[0], hx = p[1], ly = p[2], hy = p[3], lz = p[4], hz = p[5];
        const uint4   ch = reinterpret_cast<const uint4 *>(p)[6];
        const uint32_t nc = reinterpret_cast<const uint4 *>(p)[7].x;
        ref[0] = ch.x;
        ref[1] = ch.y;
        ref[2] = ch.z;
        ref[3] = ch.w;
        const float LX[4] = {lx.x, lx.y, lx.z, lx.w}, HX[4] = {hx.x, hx.y, hx.z, hx.w};
        const float LY[4] = {ly.x, ly.y, ly.z, ly.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w};
        const float LZ[4] = {lz.x, lz.y, lz.z, lz.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
        int nh = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float t;
            const bool h = (uint32_t)j < nc && slab(LX[j], HX[j], LY[j], HY[j], LZ[j], HZ[j], r, t);
            key[j] = h ? make_key(t, j) : kMissKey;
            nh += h ? 1 : 0;
        }
        return nh;
    }
};

template <>
struct NodeOps<8> {
    // Plane distances in FMA form: t = q * (2^e * inv) + (origin - o) * inv, one rounding per
    // plane instead of decoding the plane first.  The rounding of the per-node base and of
    // the FMA is bounded by delta = 2^-20 * (|base| + 255 |2^e inv|) per axis; subtracting
    // it from the base of the entry plane and adding it to the base of the exit plane keeps
    // every child interval a superset of the exact one, at no per-child cost.
    // The entry plane of an axis is the low one for a positive direction and the high one for
    // a negative direction, so the byte words are picked once per node (lo <= hi holds for
    // every child slot below the child count) and a child costs no min/max per axis.
    static __device__ __forceinline__ int expand(const void *nodes, uint32_t idx, const RayState &r, uint32_t (&key)[8],
                                                 uint32_t (&ref)[8])
    {
        const uint4 *p  = reinterpret_cast<const uint4 *>(static_cast<const Node8Q *>(nodes) + idx);
        const uint4  h  = p[0];
        const uint4  c0 = p[1], c1 = p[2];
        const uint4  qx = p[3], qy = p[4], qz = p[5];
        ref[0] = c0.x; ref[1] = c0.y; ref[2] = c0.z; ref[3] = c0.w;
        ref[4] = c1.x; ref[5] = c1.y; ref[6] = c1.z; ref[7] = c1.w;
        const float sx = __uint_as_float((h.w & 0xFFu) << 23) * r.inv.x;
        const float sy = __uint_as_float(((h.w >> 8) & 0xFFu) << 23) * r.inv.y;
        const float sz = __uint_as_float(((h.w >> 16) & 0xFFu) << 23) * r.inv.z;
        const float bx = (__uint_as_float(h.x) - r.o.x) * r.inv.x;
        const float by = (__uint_as_float(h.y) - r.o.y) * r.inv.y;
        const float bz = (__uint_as_float(h.z) - r.o.z) * r.inv.z;
        const float dx = 9.5367431640625e-07f * (fabsf(bx) + 255.0f * fabsf(sx));
        const float dy = 9.5367431640625e-07f * (fabsf(by) + 255.0f * fabsf(sy));
        const float dz = 9.5367431640625e-07f * (fabsf(bz) + 255.0f * fabsf(sz));
        const float bnx = bx - dx, bfx = bx + dx, bny = by - dy, bfy = by + dy, bnz = bz - dz, bfz = bz + dz;
        // words [0..1] hold the low bytes of children 0-3 / 4-7, words [2..3] the high bytes
        const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
        const uint32_t nx0 = px ? qx.x : qx.z, nx1 = px ? qx.y : qx.w, fx0 = px ? qx.z : qx.x, fx1 = px ? qx.w : qx.y;
        const uint32_t ny0 = py ? qy.x : qy.z, ny1 = py ? qy.y : qy.w, fy0 = py ? qy.z : qy.x, fy1 = py ? qy.w : qy.y;
        const uint32_t nz0 = pz ? qz.x : qz.z, nz1 = pz ? qz.y : qz.w, fz0 = pz ? qz.z : qz.x, fz1 = pz ? qz.w : qz.y;
        const float tn0 = fmaxf(r.tn, 0.0f);
        const uint32_t nc = h.w >> 24;
        int nh = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int sh = 8 * (j & 3);
            const float tnx = fmaf((float)(((j < 4 ? nx0 : nx1) >> sh) & 0xFFu), sx, bnx);
            const float tfx = fmaf((float)(((j < 4 ? fx0 : fx1) >> sh) & 0xFFu), sx, bfx);
            const float tny = fmaf((float)(((j < 4 ? ny0 : ny1) >> sh) & 0xFFu), sy, bny);
            const float tfy = fmaf((float)(((j < 4 ? fy0 : fy1) >> sh) & 0xFFu), sy, bfy);
            const float tnz = fmaf((float)(((j < 4 ? nz0 : nz1) >> sh) & 0xFFu), sz, bnz);
            const float tfz = fmaf((float)(((j < 4 ? fz0 : fz1) >> sh) & 0xFFu), sz, bfz);
            const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tn0));
            const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, r.tmax));
            const bool hit = (uint32_t)j < nc && tmin <= tmax;
            key[j] = hit ? make_key(tmin, j) : kMissKey;
            nh += hit ? 1 : 0;
        }
        return nh;
    }
};

template <uint32_t S>
struct Stack {
    uint2    *s;        // LDS [S + 1][kTraceBlock] of (ref, entry distance bits); row S absorbs
                        // the writes of children that missed
    uint4    *g;        // this lane's global spill: kSpillDepth (ref, distance) entries, contiguous, so
                        // that a spill of S entries writes S * 8 bytes of whole lines (uint4 = 2 entries)
    uint32_t  tid;
    uint32_t  sp, gsp;

    uint32_t *err;

    __device__ __forceinline__ void push(uint32_t ref, float dist)
    {
        if (sp == S) {
            if (gsp + S > kSpillDepth) {
                // cannot happen for trees admitted at commit; never write out of bounds
                *err = 1u;
                return;
            }
            for (uint32_t k = 0; k < S; k += 2) {
                const uint2 a = s[k * kTraceBlock + tid], b = s[(k + 1) * kTraceBlock + tid];
                g[(gsp + k) / 2] = make_uint4(a.x, a.y, b.x, b.y);
            }
            gsp += S;
            sp = 0;
        }
        s[sp * kTraceBlock + tid] = make_uint2(ref, __float_as_uint(dist));
        ++sp;
    }
    // Pushes the hit children of a node, keys sorted ascending with the nh - 1 hits behind
    // the nearest in slots 1..nh-1, so that slot 1 ends on top.  When they fit, every slot
    // is written unconditionally (misses into row S), which keeps the node step free of
    // per-child branches.
    template <int W>
    __device__ __forceinline__ void push_children(const uint32_t (&key)[W], const uint32_t (&ref)[W], int nh)
    {
        const uint32_t m = (uint32_t)nh - 1u;
        if (sp + m <= S) {
#pragma unroll
            for (int k = 1; k < W; ++k) {
                const uint32_t pos = k < nh ? sp + m - (uint32_t)k : S;
                s[pos * kTraceBlock + tid] = make_uint2(ref[k], key[k] & ~7u);
            }
            sp += m;
        } else {
#pragma unroll
            for (int k = W - 1; k >= 1; --k) {
                if (key[k] != kMissKey) push(ref[k], __uint_as_float(key[k] & ~7u));
            }
        }
    }
    // pops the next entry closer than tmax; kEmpty when the stack is exhausted
    __device__ __forceinline__ uint32_t pop(float tmax)
    {
        for (;;) {
            if (sp == 0) {
                if (gsp == 0) return kEmpty;
                gsp -= S;
                for (uint32_t k = 0; k < S; k += 2) {
                    const uint4 v = g[(gsp + k) / 2];
                    s[k * kTraceBlock + tid]       = make_uint2(v.x, v.y);
                    s[(k + 1) * kTraceBlock + tid] = make_uint2(v.z, v.w);
                }
                sp = S;
            }
            --sp;
            const uint2 v = s[sp * kTraceBlock + tid];
            if (__uint_as_float(v.y) <= tmax) return v.x;
        }
    }
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// kMulti: rays of many scenes in one launch (batched rendering); each lane takes its
// ray's scene (nodes, triangles, root) when it fetches the ray.
// kInst: two-level scene.  A leaf of the instance BVH (level 0) names instances; entering
// one transforms the lane's ray into the instance's object space (direction renormalised,
// distances scaled by |inv * d|, model.cpp:107-113) and pushes kInstExit, whose pop
// restores the world ray; closest-hit distances are compared in world units.
// The 8-entry stack variant is also the high-occupancy one: its 18 KB of LDS per block
// leaves room for 8 waves per SIMD, so the register budget is capped to match (64 VGPRs).
// kStats: the instrumented variant; counts each lane's inner-node tests, leaf visits and
// triangle tests, each wave's iterations of the node and leaf loops, and the lanes idle in
// node-loop iterations by cause (no ray, ray finished but for parked leaves, waiting at a
// leaf with both parking slots full), and adds them to q.tstats when the wave leaves.
// The launch's epilogue (TraceQueue::done): every block reports when its waves are done with
// the work counters; the last one clears the next bounce's counters.  No fence is needed:
// every work-counter atomic returned (its value was used) before the block's barrier, the
// last block reads nothing other blocks wrote, and the end of the kernel publishes its
// stores to the next launch.  (A device-scope fence here would write back and invalidate
// the XCD's L2 under the blocks still traversing.)
__device__ __forceinline__ void trace_epilogue(const TraceQueue &q, uint32_t n)
{
    if (!q.done) return;
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (atomicAdd(q.done, 1u) != gridDim.x - 1u) return;
    for (uint32_t x = 0; x < kXcds; ++x) q.counters[x] = 0u;
    if (q.clear0) *q.clear0 = 0u;
    if (q.clear1) *q.clear1 = 0u;
    if (q.stat) atomicAdd(q.stat, (unsigned long long)n);
    *q.done = 0u;
}

template <int W, bool kAny, bool kMulti, uint32_t S, bool kInst, bool kStats = false>
__global__ __launch_bounds__(kTraceBlock, S == 8 ? 8 : 1) void k_trace(SceneDev sc, TraceQueue q, uint8_t *occ,
                                                                      uint32_t *spill)
{
    uint32_t c_nodes = 0, c_leaves = 0, c_tris = 0, c_wnode = 0, c_wleaf = 0;
    uint32_t c_noray = 0, c_empty = 0, c_leafwait = 0;   // idle lanes of node-loop iterations, by cause
    __shared__ uint2 s_stack[(S + 1) * kTraceBlock];
    const uint32_t n = *q.count;
    if (n == 0) {
        trace_epilogue(q, n);
        return;
    }
    const uint32_t lane  = lane_id();
    const uint32_t xcd   = blockIdx.x % kXcds;
    const uint32_t slice = (n + kXcds - 1) / kXcds;

    Stack<S> st;
    st.s      = s_stack;
    st.g      = reinterpret_cast<uint4 *>(spill) + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * (kSpillDepth / 2);
    st.tid    = threadIdx.x;
    st.sp = st.gsp = 0;
    st.err = q.error;

    RayState r;
    const void   *nodes = sc.nodes;
    const float4 *tris  = sc.tri;
    uint32_t rid = 0, cur = kEmpty;
    // instance state (kInst): world ray while inside an instance, scale, instance ids
    f3       wo = mk3(0, 0, 0), wd = mk3(0, 0, 0);
    float    wtmax = 0.0f, ilen = 1.0f;
    uint32_t level = 0, cur_inst = kInvalidRef, best_inst = kInvalidRef;
    bool     inst_hit = false;
    auto enter_instance = [&](uint32_t inst) {
        const InstDev &I = sc.inst[inst];
        wo    = r.o;
        wd    = r.d;
        wtmax = r.tmax;
        const f3 o2 = mk3(I.inv[0] * r.o.x + I.inv[1] * r.o.y + I.inv[2] * r.o.z + I.inv[3],
                          I.inv[4] * r.o.x + I.inv[5] * r.o.y + I.inv[6] * r.o.z + I.inv[7],
                          I.inv[8] * r.o.x + I.inv[9] * r.o.y + I.inv[10] * r.o.z + I.inv[11]);
        const f3 d2 = mk3(I.inv[0] * r.d.x + I.inv[1] * r.d.y + I.inv[2] * r.d.z,
                          I.inv[4] * r.d.x + I.inv[5] * r.d.y + I.inv[6] * r.d.z,
                          I.inv[8] * r.d.x + I.inv[9] * r.d.y + I.inv[10] * r.d.z);
        ilen     = sqrtf(dot(d2, d2));
        r.o      = o2;
        r.d      = d2 * (1.0f / ilen);
        r.inv    = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
        r.wt     = wt_setup(r.d);
        r.tmax   = wtmax * ilen;
        level    = 1;
        cur_inst = inst;
        inst_hit = false;
        st.push(kInstExit, -kInf);
        return I.root;
    };
    auto exit_instance = [&]() {
        r.o    = wo;
        r.d    = wd;
        r.inv  = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
        r.wt   = wt_setup(r.d);
        r.tmax = inst_hit ? r.tmax / ilen : wtmax;
        level  = 0;
    };
    // next node to visit: pops, leaving instances as their exit markers come up
    auto next = [&]() -> uint32_t {
        for (;;) {
            const uint32_t c = st.pop(r.tmax);
            if (!kInst || c != kInstExit) return c;
            exit_instance();
        }
    };
    // leaves parked by the inner-node phase, pend first (not for two-level scenes, whose
    // instance leaves change the lane's ray)
    constexpr bool kSpec = !kInst;
    uint32_t pend = kEmpty, pend2 = kEmpty;
    bool     active = false, exhausted = false;
    uint32_t next_slice = 0;   // slices tried by this wave (wave-uniform)

    for (;;) {
        // ---- refill idle lanes ----
        uint64_t idle = __ballot(!active);
        uint32_t nidle = __popcll(idle);
        while (!exhausted && nidle >= q.refill_idle) {
            uint32_t base = 0, got = 0;
            if (lane == 0) {
                while (next_slice < kXcds) {
                    const uint32_t x   = (xcd + next_slice) % kXcds;
                    const uint32_t beg = x * slice;
                    const uint32_t len = beg < n ? min(slice, n - beg) : 0u;
                    const uint32_t b   = len ? atomicAdd(q.counters + x, nidle) : len;
                    if (b < len) {
                        base = beg + b;
                        got  = min(nidle, len - b);
                        break;
                    }
                    ++next_slice;
                }
            }
            base       = __shfl(base, 0);
            got        = __shfl(got, 0);
            next_slice = __shfl(next_slice, 0);
            if (got == 0) {
                exhausted = true;
                break;
            }
            const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
            if (!active && rank < got) {
                rid = base + rank;
                r.o   = mk3(q.rays.ox[rid], q.rays.oy[rid], q.rays.oz[rid]);
                r.d   = mk3(q.rays.dx[rid], q.rays.dy[rid], q.rays.dz[rid]);
                if (q.unit_dir && (r.d.x != 0.0f || r.d.y != 0.0f || r.d.z != 0.0f)) r.d = normalize(r.d);
                r.tn  = q.rays.tn[rid];
                r.tmax = q.rays.tf[rid];
                r.inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
                r.wt  = wt_setup(r.d);
                r.best_tri = kInvalidRef;
                r.best_prim = kInvalidRef;
                r.bu = r.bv = 0.0f;
                st.sp = st.gsp = 0;
                pend      = kEmpty;
                pend2     = kEmpty;
                level     = 0;
                best_inst = kInvalidRef;
                uint32_t root = sc.root, ntri = sc.ntri;
                if constexpr (kMulti) {
                    const TraceScene ts = q.scenes[q.path_job[q.ray_path[rid]]];
                    nodes = ts.nodes;
                    tris  = ts.tri;
                    root  = ts.root;
                    ntri  = ts.ntri;
                }
                const bool valid = ntri > 0 && r.tn <= r.tmax && r.tmax == r.tmax && r.tn == r.tn &&
                                   (r.d.x != 0.0f || r.d.y != 0.0f || r.d.z != 0.0f);
                cur    = valid ? root : kEmpty;
                active = true;
            }
            idle  = __ballot(!active);
            nidle = __popcll(idle);
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        // ---- inner-node phase: lanes expand nodes until every lane holds a leaf or is done.
        // kSpec: a lane whose next node is a leaf parks it (two slots) and keeps traversing
        // until its slots are full, so that the leaf phase runs with more lanes busy ----
        for (;;) {
            if (kSpec && active && pend2 == kEmpty && cur != kEmpty && ref_is_leaf(cur)) {
                if (pend == kEmpty) pend = cur;
                else pend2 = cur;
                cur = next();
            }
            const bool inner = active && cur != kEmpty && !ref_is_leaf(cur);
            const uint64_t srch = __ballot(inner && (!kSpec || pend2 == kEmpty));
            if (srch == 0) break;
            // fewer than park_min lanes still searching and some holding a leaf: run the leaf
            // phase now rather than keep most of the wave waiting
            if (kSpec && __popcll(srch) < q.park_min && __ballot(active && pend != kEmpty) != 0) break;
            if constexpr (kStats) {
                ++c_wnode;
                c_noray += !active;
                c_empty += active && cur == kEmpty;
                c_leafwait += active && cur != kEmpty && ref_is_leaf(cur);
            }
            if (inner) {
                uint32_t key[W], ref[W];
                const int nh = NodeOps<W>::expand(nodes, cur, r, key, ref);
                if constexpr (kStats) ++c_nodes;
                if (nh == 0) {
                    cur = next();
                } else {
                    sort_pairs<W>(key, ref);
                    cur = ref[0];
                    st.template push_children<W>(key, ref, nh);
                }
            }
        }
        // ---- leaf phase: the parked leaf first, then leaves popped after it ----
        for (uint32_t round = 0;; ++round) {
            const bool parked = kSpec && active && pend != kEmpty;
            const bool leaf   = active && (parked || (cur != kEmpty && ref_is_leaf(cur)));
            if (__ballot(leaf) == 0) break;
            // a lane left holding an unparked leaf parks it in the next inner-node phase
            if (kSpec && q.leaf_rounds && round >= q.leaf_rounds && __ballot(parked) == 0) break;
            if constexpr (kStats) ++c_wleaf;
            if (kInst && leaf && level == 0) {
                // instance leaf: enter the first instance, keep the rest for later
                const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
                if (cnt > 1) st.push(ref_leaf(first + 1, cnt - 1), -kInf);
                cur = enter_instance(sc.tlas_inst[first]);
            } else if (leaf) {
                const uint32_t lf = parked ? pend : cur;
                const uint32_t first = leaf_first(lf), cnt = leaf_count(lf);
                bool any = false;
                if constexpr (kStats) ++c_leaves;
                for (uint32_t k = 0; k < cnt; ++k) {
                    if constexpr (kStats) ++c_tris;
                    const size_t  ti = (size_t)(first + k);
                    const float4  a  = tris[3 * ti + 0];
                    const float4  b  = tris[3 * ti + 1];
                    const float4  c  = tris[3 * ti + 2];
                    float t, u, v;
                    if (wt_test(r.wt, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), r.o, r.tn, r.tmax,
                                t, u, v)) {
                        const uint32_t prim = __float_as_uint(a.w);
                        if (t < r.tmax || prim < r.best_prim) {
                            r.tmax      = t;
                            r.best_tri  = (uint32_t)ti;
                            r.best_prim = prim;
                            r.bu        = u;
                            r.bv        = v;
                            if (kInst) {
                                best_inst = cur_inst;
                                inst_hit  = true;
                            }
                        }
                        if (kAny) {
                            any = true;
                            break;
                        }
                    }
                }
                if (parked) {
                    pend  = pend2;
                    pend2 = kEmpty;
                    if (kAny && any) {
                        cur  = kEmpty;
                        pend = kEmpty;
                    }
                } else {
                    cur = (kAny && any) ? kEmpty : next();
                }
            }
        }
        if (active && cur == kEmpty) {
            if (kAny) {
                occ[rid] = r.best_tri != kInvalidRef ? 1 : 0;
            } else {
                const bool h = r.best_tri != kInvalidRef;
                q.hits.t[rid]   = h ? r.tmax : kInf;
                q.hits.tri[rid] = r.best_tri;
                if (q.hits.u) q.hits.u[rid] = r.bu;
                if (q.hits.v) q.hits.v[rid] = r.bv;
                if (kInst && q.hits.inst) q.hits.inst[rid] = h ? best_inst : kInvalidRef;
            }
            active = false;
        }
    }
    if constexpr (kStats) {
        unsigned long long a = c_nodes, b = c_leaves, c = c_tris, d = c_noray, e = c_empty, f = c_leafwait;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
            c += __shfl_xor(c, o);
            d += __shfl_xor(d, o);
            e += __shfl_xor(e, o);
            f += __shfl_xor(f, o);
        }
        if (lane == 0 && q.tstats) {
            atomicAdd(q.tstats + 0, a);
            atomicAdd(q.tstats + 1, b);
            atomicAdd(q.tstats + 2, c);
            atomicAdd(q.tstats + 3, (unsigned long long)c_wnode);   // wave-uniform counts
            atomicAdd(q.tstats + 4, (unsigned long long)c_wleaf);
            atomicAdd(q.tstats + 5, d);
            atomicAdd(q.tstats + 6, e);
            atomicAdd(q.tstats + 7, f);
        }
    }
    trace_epilogue(q, n);
}

}  // namespace

uint32_t trace_grid(int device, uint32_t blocks_per_cu)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return 1024;
    // two 256-lane blocks per CU: fewer rays in flight keep more of the BVH in cache
    if (blocks_per_cu == 0) blocks_per_cu = 2;
    uint32_t g = (uint32_t)p.multiProcessorCount * blocks_per_cu;
    // whole multiple of the XCD count so every XCD gets the same number of blocks
    g = (g + kXcds - 1) / kXcds * kXcds;
    return g;
}

uint64_t trace_spill_bytes(uint32_t grid)
{
    return (uint64_t)grid * kTraceBlock * kSpillDepth * 2u * sizeof(uint32_t);
}

namespace {

// Instantiates k_trace for the BVH width and LDS stack depth of the launch.
template <bool kAny, bool kMulti, bool kInst>
hipError_t dispatch_trace(uint32_t width, const TraceConfig &tc, const SceneDev &sc, const TraceQueue &qin, uint8_t *occ,
                          uint32_t *spill, hipStream_t s)
{
    const dim3 g(tc.grid), b(kTraceBlock);
    TraceQueue q   = qin;
    q.refill_idle  = tc.refill ? std::min(tc.refill, 64u) : kRefillIdle;
    q.park_min     = tc.park_min ? std::min(tc.park_min, 64u) : kParkMin;
    q.leaf_rounds  = tc.leaf_rounds ? (tc.leaf_rounds > 64u ? 0u : tc.leaf_rounds) : kLeafRounds;
    if (!kMulti && q.tstats) {
        // instrumented traversal: one stack depth (the counts do not depend on it)
        if (width == 4) k_trace<4, kAny, false, 32, kInst, true><<<g, b, 0, s>>>(sc, q, occ, spill);
        else k_trace<8, kAny, false, 32, kInst, true><<<g, b, 0, s>>>(sc, q, occ, spill);
        return hipGetLastError();
    }
#define CRT_TRACE(Wd, Sd) k_trace<Wd, kAny, kMulti, Sd, kInst><<<g, b, 0, s>>>(sc, q, occ, spill)
    if (width == 4) {
        if (tc.stack == 8) CRT_TRACE(4, 8);
        else if (tc.stack == 32) CRT_TRACE(4, 32);
        else CRT_TRACE(4, 16);
    } else {
        if (tc.stack == 8) CRT_TRACE(8, 8);
        else if (tc.stack == 32) CRT_TRACE(8, 32);
        else CRT_TRACE(8, 16);
    }
#undef CRT_TRACE
    return hipGetLastError();
}

}  // namespace

hipError_t launch_trace_closest(const SceneDev &sc, const TraceQueue &q, uint32_t *spill, const TraceConfig &tc,
                                hipStream_t s)
{
    if (sc.instanced) return dispatch_trace<false, false, true>(sc.width, tc, sc, q, nullptr, spill, s);
    return dispatch_trace<false, false, false>(sc.width, tc, sc, q, nullptr, spill, s);
}

hipError_t launch_trace_any(const SceneDev &sc, const TraceQueue &q, uint8_t *occ, uint32_t *spill,
                            const TraceConfig &tc, hipStream_t s)
{
    if (sc.instanced) return dispatch_trace<true, false, true>(sc.width, tc, sc, q, occ, spill, s);
    return dispatch_trace<true, false, false>(sc.width, tc, sc, q, occ, spill, s);
}

hipError_t launch_trace_closest_multi(uint32_t width, const TraceQueue &q, uint32_t *spill, const TraceConfig &tc,
                                      hipStream_t s)
{
    SceneDev none{};
    none.ntri = 1;   // per-ray scenes decide
    return dispatch_trace<false, true, false>(width, tc, none, q, nullptr, spill, s);
}

hipError_t launch_trace_any_multi(uint32_t width, const TraceQueue &q, uint8_t *occ, uint32_t *spill,
                                  const TraceConfig &tc, hipStream_t s)
{
    SceneDev none{};
    none.ntri = 1;
    return dispatch_trace<true, true, false>(width, tc, none, q, occ, spill, s);
}

}  // namespace crt
